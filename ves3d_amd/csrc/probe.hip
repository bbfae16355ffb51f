// probe.hip -- the VALU issue rate this GPU sustains, as a reference for the
// far-field kernels' rooflines: independent fused multiply-add chains in
// FP32 (non-packed, as every kernel of this library is built) and FP64,
// timed with HIP events.  No counterpart in the reference.
#include "common.hpp"

namespace v3d {
namespace {

template <typename T, int CH>
__global__ __launch_bounds__(256) void k_fma_chains(T *out, int iters, T a, T b) {
  T v[CH];
#pragma unroll
  for (int c = 0; c < CH; ++c) v[c] = (T)threadIdx.x * (T)1e-3 + (T)c;
  for (int i = 0; i < iters; ++i) {
#pragma unroll
    for (int c = 0; c < CH; ++c) v[c] = fma(v[c], a, b);
  }
  T s = 0;
#pragma unroll
  for (int c = 0; c < CH; ++c) s += v[c];
  out[(long)blockIdx.x * 256 + threadIdx.x] = s;
}

template <typename T>
int run(double *rate) {
  constexpr int CH = sizeof(T) == 4 ? 16 : 8;
  const int blocks = 256 * 32, iters = 4096;
  T *out = nullptr;
  V3D_HIP(hipMalloc(&out, sizeof(T) * blocks * 256));
  hipEvent_t a, b;
  V3D_HIP(hipEventCreate(&a));
  V3D_HIP(hipEventCreate(&b));
  float best = 1e30f;
  for (int rep = 0; rep < 3; ++rep) {  // the first launch also warms the clocks
    V3D_HIP(hipEventRecord(a, nullptr));
    k_fma_chains<T, CH><<<blocks, 256>>>(out, iters, (T)0.999, (T)1e-3);
    V3D_HIP(hipEventRecord(b, nullptr));
    V3D_HIP(hipEventSynchronize(b));
    float ms = 0;
    V3D_HIP(hipEventElapsedTime(&ms, a, b));
    if (rep > 0 && ms < best) best = ms;
  }
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  (void)hipFree(out);
  *rate = (double)blocks * 256 * iters * CH / (best * 1e-3);
  return VES3D_SUCCESS;
}

}  // namespace
}  // namespace v3d

extern "C" int ves3d_hip_valu_probe(int fp64, double *lane_fma_per_s) {
  if (!lane_fma_per_s) return VES3D_INVALID_PARAMETER;
  return fp64 ? v3d::run<double>(lane_fma_per_s) : v3d::run<float>(lane_fma_per_s);
}

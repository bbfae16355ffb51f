// common.hpp -- shared definitions for the gfx950 HIP kernels of ves3d_amd.
#pragma once
#include <hip/hip_runtime.h>
#include <cstddef>
#include <cstdint>
#include <cstdio>

#include "../../include/ves3d_hip.h"

namespace v3d {

constexpr double kPi = 3.14159265358979323846;

// Launch-error check used after every kernel launch: converts a HIP error into
// the reference's Error_t numbering (include/Error.h:46-67).
inline int hip_status(hipError_t e) {
  if (e == hipSuccess) return VES3D_SUCCESS;
  fprintf(stderr, "[ves3d_hip] HIP error: %s\n", hipGetErrorString(e));
  if (e == hipErrorOutOfMemory) return VES3D_MEMORY_ERROR;
  if (e == hipErrorInvalidValue) return VES3D_INVALID_PARAMETER;
  if (e == hipErrorInvalidDevice || e == hipErrorNoDevice) return VES3D_INVALID_DEVICE;
  return VES3D_UNKNOWN_ERROR;
}

inline int launch_status() { return hip_status(hipGetLastError()); }

#define V3D_CHECK(expr)                         \
  do {                                          \
    int _s = (expr);                            \
    if (_s != VES3D_SUCCESS) return _s;         \
  } while (0)

#define V3D_HIP(expr) V3D_CHECK(::v3d::hip_status(expr))

inline unsigned cdiv(size_t a, size_t b) { return (unsigned)((a + b - 1) / b); }

// Wave64 sum reduction through DPP/cross-lane shuffles.
template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Error-free transformation a+b = s+e (Knuth TwoSum), used for compensated
// FP32 accumulation of per-tile partial sums.
__device__ __forceinline__ void two_sum(float a, float b, float &s, float &e) {
  s = a + b;
  float bb = s - a;
  e = (a - (s - bb)) + (b - bb);
}

}  // namespace v3d

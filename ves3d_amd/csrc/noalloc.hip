// noalloc.hip -- an allocation-free form of the far-field evaluation
// (ves3d_hip_direct_eval_f64, the PVFMMEval stand-in) for callers that cannot
// call hipMalloc: code under HIP graph capture, tight serving loops,
// harnesses that forbid allocation.  Its one scratch buffer, the per-slice
// FP64 partial sums of the direct sum, is the fixed device array `scratch`
// below (64 MiB on every GPU the library runs on, held until the process
// ends) instead of the library's growable workspace.
#include <mutex>

#include "common.hpp"
#include "implicit.hpp"

namespace v3d {

constexpr size_t kNoAllocScratchBytes = size_t(64) << 20;

static __device__ __align__(256) double scratch[kNoAllocScratchBytes / sizeof(double)];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_mu;
double *g_scratch[kMaxDevices] = {};  // device address of `scratch`, per device

// The array's address on the current device, resolved once per device (each
// GPU holds its own copy at its own address).  `scratch` itself is a
// host-side stand-in and must never be handed to a kernel or a copy.
int scratch_on_current_device(double **out) {
  int dev = 0;
  V3D_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return VES3D_INVALID_DEVICE;
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_scratch[dev]) {
    double *p = nullptr;
    const hipError_t e = hipGetSymbolAddress((void **)&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) return VES3D_UNKNOWN_ERROR;
    g_scratch[dev] = p;
  }
  *out = g_scratch[dev];
  return VES3D_SUCCESS;
}
}  // namespace

}  // namespace v3d

using namespace v3d;

extern "C" int ves3d_hip_direct_eval_f64_noalloc(const double *src_pos, const double *sl_den,
                                                 const double *dl_den, size_t n_src,
                                                 const double *trg_pos, double *trg_vel,
                                                 size_t n_trg, size_t spg, size_t tpg,
                                                 ves3d_stream_t stream) {
  if (n_trg > kNoAllocScratchBytes / (3 * sizeof(double))) {
    fprintf(stderr,
            "[ves3d_hip] direct_eval_f64_noalloc: %zu targets exceed the fixed scratch "
            "(%zu targets at most)\n",
            n_trg, kNoAllocScratchBytes / (3 * sizeof(double)));
    return VES3D_INVALID_PARAMETER;
  }
  double *part = nullptr;
  V3D_CHECK(scratch_on_current_device(&part));
  return direct_eval_f64_fixed(src_pos, sl_den, dl_den, n_src, trg_pos, trg_vel, n_trg, spg, tpg,
                               part, kNoAllocScratchBytes, (hipStream_t)stream);
}

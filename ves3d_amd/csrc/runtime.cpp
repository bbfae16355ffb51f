// runtime.cpp -- Device<GPU> constructor / memory surface (include/Device.h:95-119)
// and the library's cached workspaces.
#include <hip/hip_runtime.h>
#include <cstring>
#include <mutex>

#include "common.hpp"
#include "runtime.hpp"

namespace v3d {

namespace {
struct Slot {
  void *ptr = nullptr;
  size_t bytes = 0;
};
Slot g_slots[64][kWsCount];
std::mutex g_mu;
}  // namespace

void *workspace(int slot, size_t bytes) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return nullptr;
  std::lock_guard<std::mutex> lk(g_mu);
  Slot &s = g_slots[dev][slot];
  if (s.bytes >= bytes && s.ptr) return s.ptr;
  if (s.ptr) {
    (void)hipDeviceSynchronize();
    (void)hipFree(s.ptr);
    s.ptr = nullptr;
    s.bytes = 0;
  }
  size_t want = bytes + bytes / 4 + 256;
  if (hipMalloc(&s.ptr, want) != hipSuccess) {
    s.ptr = nullptr;
    return nullptr;
  }
  s.bytes = want;
  return s.ptr;
}

}  // namespace v3d

using namespace v3d;

extern "C" {

int ves3d_hip_init(int device_id) {
  int n = 0;
  V3D_HIP(hipGetDeviceCount(&n));
  if (device_id < 0 || device_id >= n) return VES3D_INVALID_DEVICE;
  V3D_HIP(hipSetDevice(device_id));
  hipDeviceProp_t prop;
  V3D_HIP(hipGetDeviceProperties(&prop, device_id));
  if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) {
    fprintf(stderr, "[ves3d_hip] device %d is %s, this library is built for gfx950\n",
            device_id, prop.gcnArchName);
    return VES3D_INVALID_DEVICE;
  }
  return VES3D_SUCCESS;
}

int ves3d_hip_device_info(char *name_host, size_t name_len, int *num_cu_host,
                          int *clock_mhz_host) {
  int dev = 0;
  V3D_HIP(hipGetDevice(&dev));
  hipDeviceProp_t prop;
  V3D_HIP(hipGetDeviceProperties(&prop, dev));
  if (name_host && name_len) {
    std::strncpy(name_host, prop.name, name_len - 1);
    name_host[name_len - 1] = 0;
  }
  if (num_cu_host) *num_cu_host = prop.multiProcessorCount;
  if (clock_mhz_host) *clock_mhz_host = prop.clockRate / 1000;
  return VES3D_SUCCESS;
}

int ves3d_hip_malloc(void **ptr_host, size_t nbytes) {
  if (!ptr_host) return VES3D_INVALID_PARAMETER;
  return hip_status(hipMalloc(ptr_host, nbytes ? nbytes : 1));
}

int ves3d_hip_free(void *ptr) { return hip_status(hipFree(ptr)); }

int ves3d_hip_memcpy(void *dst, const void *src, size_t nbytes, int kind) {
  static const hipMemcpyKind kinds[4] = {hipMemcpyHostToHost, hipMemcpyHostToDevice,
                                         hipMemcpyDeviceToHost, hipMemcpyDeviceToDevice};
  if (kind < 0 || kind > 3) return VES3D_INVALID_PARAMETER;
  if (nbytes == 0) return VES3D_SUCCESS;
  return hip_status(hipMemcpy(dst, src, nbytes, kinds[kind]));
}

int ves3d_hip_memset(void *ptr, int value, size_t nbytes) {
  if (nbytes == 0) return VES3D_SUCCESS;
  return hip_status(hipMemset(ptr, value, nbytes));
}

int ves3d_hip_synchronize(void) { return hip_status(hipDeviceSynchronize()); }

}  // extern "C"

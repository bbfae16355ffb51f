// vecops.hip -- BLAS-1 kernels.  The dot products replace the reference's
// AlgebraicDot (src/DeviceCPU.cc, src/CudaKernels.cu:1114-1188, whose GPU
// form only uses block 0) with a two-stage, fixed-order FP64 reduction.
#include <algorithm>
#include <cmath>

#include "runtime.hpp"
#include "vecops.hpp"

namespace v3d {

namespace {
constexpr int RB = 256;      // threads per block
constexpr int RBLOCKS = 512; // first-stage blocks

// x may alias y, so neither pointer is restrict-qualified
__global__ void k_axpby(long n, double a, const double *x, double b, double *y) {
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    y[i] = a * x[i] + b * y[i];
}
// y = a x without reading y (b == 0: y may hold uninitialised memory, whose
// NaN bit patterns 0 * y would propagate)
__global__ void k_scale(long n, double a, const double *x, double *y) {
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    y[i] = a * x[i];
}
__global__ void k_axpy3(long n, double a, const double *__restrict__ x,
                        const double *__restrict__ y, double *__restrict__ z) {
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    z[i] = a * x[i] + y[i];
}
__global__ void k_fill(long n, double v, double *__restrict__ y) {
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    y[i] = v;
}

struct DotArgs {
  const double *x[4];
  const double *y[4];
};

__global__ __launch_bounds__(RB) void k_dots_stage1(long n, int nd, DotArgs a,
                                                     double *__restrict__ part) {
  __shared__ double sh[4][RB / 64];
  double acc[4] = {0, 0, 0, 0};
  for (long i = (long)blockIdx.x * RB + threadIdx.x; i < n; i += (long)gridDim.x * RB)
#pragma unroll
    for (int k = 0; k < 4; ++k)
      if (k < nd) acc[k] += a.x[k][i] * a.y[k][i];
  const int lane = threadIdx.x % 64, w = threadIdx.x / 64;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    double v = wave_sum(acc[k]);
    if (lane == 0) sh[k][w] = v;
  }
  __syncthreads();
  if (threadIdx.x < nd) {
    double s = 0;
    for (int j = 0; j < RB / 64; ++j) s += sh[threadIdx.x][j];
    part[threadIdx.x * RBLOCKS + blockIdx.x] = s;
  }
}

__global__ __launch_bounds__(RB) void k_dots_stage2(int nd, int nb, const double *__restrict__ part,
                                                     double *__restrict__ out) {
  __shared__ double sh[RB / 64];
  for (int k = 0; k < nd; ++k) {
    double s = 0;
    for (int j = threadIdx.x; j < nb; j += RB) s += part[k * RBLOCKS + j];
    s = wave_sum(s);
    if (threadIdx.x % 64 == 0) sh[threadIdx.x / 64] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
      double t = 0;
      for (int j = 0; j < RB / 64; ++j) t += sh[j];
      out[k] = t;
    }
    __syncthreads();
  }
}

// Dots of nk vectors against one: part[k][b] = sum over block b's share of
// V[k][i] w[i]; workgroup row k of the grid serves vector k, and w is re-read
// (from cache) per vector.
__global__ __launch_bounds__(RB) void k_mdot_stage1(long n, const double *__restrict__ V,
                                                     long ldv, const double *__restrict__ w,
                                                     double *__restrict__ part) {
  __shared__ double sh[RB / 64];
  const double *v = V + (long)blockIdx.y * ldv;
  double acc = 0;
  for (long i = (long)blockIdx.x * RB + threadIdx.x; i < n; i += (long)gridDim.x * RB)
    acc += v[i] * w[i];
  acc = wave_sum(acc);
  if (threadIdx.x % 64 == 0) sh[threadIdx.x / 64] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = 0;
    for (int j = 0; j < RB / 64; ++j) s += sh[j];
    part[(long)blockIdx.y * RBLOCKS + blockIdx.x] = s;
  }
}

__global__ __launch_bounds__(RB) void k_mdot_stage2(int nb, const double *__restrict__ part,
                                                     double *__restrict__ out) {
  __shared__ double sh[RB / 64];
  double s = 0;
  for (int j = threadIdx.x; j < nb; j += RB) s += part[(long)blockIdx.x * RBLOCKS + j];
  s = wave_sum(s);
  if (threadIdx.x % 64 == 0) sh[threadIdx.x / 64] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0;
    for (int j = 0; j < RB / 64; ++j) t += sh[j];
    out[blockIdx.x] = t;
  }
}

// w = b w + sum_k c[k] V[k], the coefficients passed by value
__global__ void k_combine(long n, int nk, const double *__restrict__ V, long ldv, VecCoefs c,
                          double b, double *__restrict__ w) {
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long)gridDim.x * blockDim.x) {
    double s = b == 0.0 ? 0.0 : b * w[i];
    for (int k = 0; k < nk; ++k) s += c.c[k] * V[k * ldv + i];
    w[i] = s;
  }
}

unsigned blocks_for(long n) {
  long b = (n + RB - 1) / RB;
  if (b > 4096) b = 4096;
  return (unsigned)(b < 1 ? 1 : b);
}
}  // namespace

int dev_axpby(long n, double a, const double *x, double b, double *y, hipStream_t st) {
  if (n <= 0) return VES3D_SUCCESS;
  if (b == 0.0)
    k_scale<<<blocks_for(n), RB, 0, st>>>(n, a, x, y);
  else
    k_axpby<<<blocks_for(n), RB, 0, st>>>(n, a, x, b, y);
  return launch_status();
}
int dev_axpy3(long n, double a, const double *x, const double *y, double *z, hipStream_t st) {
  if (n <= 0) return VES3D_SUCCESS;
  k_axpy3<<<blocks_for(n), RB, 0, st>>>(n, a, x, y, z);
  return launch_status();
}
int dev_copy(long n, const double *x, double *y, hipStream_t st) {
  if (n <= 0) return VES3D_SUCCESS;
  return hip_status(hipMemcpyAsync(y, x, n * sizeof(double), hipMemcpyDeviceToDevice, st));
}
int dev_fill(long n, double v, double *y, hipStream_t st) {
  if (n <= 0) return VES3D_SUCCESS;
  k_fill<<<blocks_for(n), RB, 0, st>>>(n, v, y);
  return launch_status();
}

int dev_dots(long n, int nd, const double *const *x, const double *const *y, double *out_host,
             hipStream_t st) {
  if (nd < 1 || nd > 4) return VES3D_INVALID_PARAMETER;
  double *ws = (double *)workspace(kWsSolver, sizeof(double) * (4 * RBLOCKS + 8));
  if (!ws) return VES3D_MEMORY_ERROR;
  DotArgs a{};
  for (int k = 0; k < nd; ++k) {
    a.x[k] = x[k];
    a.y[k] = y[k];
  }
  int nb = (int)std::min<long>(RBLOCKS, std::max<long>(1, (n + RB - 1) / RB));
  k_dots_stage1<<<nb, RB, 0, st>>>(n, nd, a, ws);
  V3D_CHECK(launch_status());
  k_dots_stage2<<<1, RB, 0, st>>>(nd, nb, ws, ws + 4 * RBLOCKS);
  V3D_CHECK(launch_status());
  V3D_HIP(hipMemcpyAsync(out_host, ws + 4 * RBLOCKS, nd * sizeof(double), hipMemcpyDeviceToHost, st));
  return hip_status(hipStreamSynchronize(st));
}

int dev_mdots(long n, int nk, const double *V, long ldv, const double *w, double *out_host,
              hipStream_t st) {
  if (nk < 1 || nk > kMaxVecCoefs || ldv < n) return VES3D_INVALID_PARAMETER;
  double *ws = (double *)workspace(kWsSolver, sizeof(double) * ((long)kMaxVecCoefs * RBLOCKS + kMaxVecCoefs));
  if (!ws) return VES3D_MEMORY_ERROR;
  const int nb = (int)std::min<long>(RBLOCKS, std::max<long>(1, (n + RB - 1) / RB));
  k_mdot_stage1<<<dim3(nb, nk), RB, 0, st>>>(n, V, ldv, w, ws);
  V3D_CHECK(launch_status());
  double *out = ws + (long)kMaxVecCoefs * RBLOCKS;
  k_mdot_stage2<<<nk, RB, 0, st>>>(nb, ws, out);
  V3D_CHECK(launch_status());
  V3D_HIP(hipMemcpyAsync(out_host, out, nk * sizeof(double), hipMemcpyDeviceToHost, st));
  return hip_status(hipStreamSynchronize(st));
}

int dev_combine(long n, int nk, const double *V, long ldv, const double *c_host, double b,
                double *w, hipStream_t st) {
  if (nk < 0 || nk > kMaxVecCoefs || ldv < n) return VES3D_INVALID_PARAMETER;
  if (n <= 0) return VES3D_SUCCESS;
  VecCoefs c{};
  for (int k = 0; k < nk; ++k) c.c[k] = c_host[k];
  k_combine<<<blocks_for(n), RB, 0, st>>>(n, nk, V, ldv, c, b, w);
  return launch_status();
}

double dev_norm(long n, const double *x, hipStream_t st, int *err) {
  const double *xs[1] = {x};
  double r = 0;
  int e = dev_dots(n, 1, xs, xs, &r, st);
  if (err) *err = e;
  return std::sqrt(r);
}

}  // namespace v3d

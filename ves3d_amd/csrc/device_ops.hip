// device_ops.hip -- the elementwise, reduction, transpose and GEMM members of
// the reference's Device<DT> surface (include/Device.h:121-232; CPU
// semantics src/DeviceCPU.cc:64-640) on AxisMajor data [vec][xyz][stride].
// All are HBM-bound single-pass kernels with grid-stride loops; the
// reductions accumulate in FP64 with a fixed order.
#include <algorithm>
#include <cmath>
#include <map>
#include <mutex>
#include <tuple>

#include "gemm.hpp"
#include "runtime.hpp"

namespace v3d {
namespace {

inline unsigned gb(long n) {
  long b = (n + 255) / 256;
  if (b > 65536) b = 65536;
  return (unsigned)(b < 1 ? 1 : b);
}
#define GL(i, n) \
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < (long)(n); i += (long)gridDim.x * blockDim.x)

template <typename T>
__global__ void k_dot3(const T *u, const T *v, long stride, long nv, T *out) {
  GL(i, nv * stride) {
    const long s = i % stride, b = (i / stride) * 3 * stride + s;
    out[i] = u[b] * v[b] + u[b + stride] * v[b + stride] + u[b + 2 * stride] * v[b + 2 * stride];
  }
}
template <typename T>
__global__ void k_cross3(const T *u, const T *v, long stride, long nv, T *w) {
  GL(i, nv * stride) {
    const long s = i % stride, b = (i / stride) * 3 * stride + s;
    const T u0 = u[b], u1 = u[b + stride], u2 = u[b + 2 * stride];
    const T v0 = v[b], v1 = v[b + stride], v2 = v[b + 2 * stride];
    w[b] = u1 * v2 - u2 * v1;
    w[b + stride] = u2 * v0 - u0 * v2;
    w[b + 2 * stride] = u0 * v1 - u1 * v0;
  }
}
template <typename T>
__global__ void k_sqrt(const T *x, long n, T *y) {
  GL(i, n) y[i] = sqrt(x[i]);
}
template <typename T>
__global__ void k_ax(const T *a, const T *x, long stride, long nv, T *y) {
  GL(i, nv * stride) y[i] = a[i % stride] * x[i];
}
template <typename T>
__global__ void k_xy(const T *x, const T *y, long n, T *z) {
  GL(i, n) z[i] = x[i] * y[i];
}
template <typename T>
__global__ void k_xyinv(const T *x, const T *y, long n, T *z) {
  GL(i, n) z[i] = (x ? x[i] : T(1)) / y[i];
}
template <typename T>
__global__ void k_uyinv(const T *u, const T *y, long stride, long nv, T *out) {
  GL(i, nv * 3 * stride) {
    const long s = i % stride, v = i / (3 * stride);
    out[i] = u[i] / y[v * stride + s];
  }
}
template <typename T>
__global__ void k_axpy(T a, const T *x, const T *y, long n, T *z) {
  GL(i, n) z[i] = y ? a * x[i] + y[i] : a * x[i];
}
template <typename T>
__global__ void k_apx(const T *a, const T *x, long stride, long ns, T *out) {
  GL(i, ns * stride) out[i] = a[i / stride] + x[i];
}
template <typename T>
__global__ void k_avpw(const T *a, const T *v, const T *w, long stride, long nv, T *out) {
  GL(i, nv * 3 * stride) {
    const T t = a[i / (3 * stride)] * v[i];
    out[i] = w ? t + w[i] : t;
  }
}
template <typename T>
__global__ void k_xvpw(const T *x, const T *v, const T *w, long stride, long nv, T *out) {
  GL(i, nv * 3 * stride) {
    const long s = i % stride, vv = i / (3 * stride);
    const T t = x[vv * stride + s] * v[i];
    out[i] = w ? t + w[i] : t;
  }
}
// Reduce (src/DeviceCPU.cc:368-410): one workgroup per (function, component)
template <typename T>
__global__ __launch_bounds__(256) void k_reduce(const T *x, int x_dim, const T *w,
                                                const T *qw, long stride, T *out) {
  const long ii = blockIdx.x / x_dim;
  const int kk = blockIdx.x % x_dim;
  double acc = 0;
  for (long j = threadIdx.x; j < stride; j += 256) {
    const double wv = (double)w[ii * stride + j] * (double)qw[j];
    acc += x ? (double)x[(ii * x_dim + kk) * stride + j] * wv : wv;
  }
  __shared__ double sh[4];
  acc = wave_sum(acc);
  if (threadIdx.x % 64 == 0) sh[threadIdx.x / 64] = acc;
  __syncthreads();
  if (threadIdx.x == 0) out[ii * x_dim + kk] = (T)(sh[0] + sh[1] + sh[2] + sh[3]);
}
// two-stage reductions: op 0 = max |x|, 1 = sum x*y, 2 = count of non-finite
template <typename T>
__global__ __launch_bounds__(256) void k_reduce1(const T *x, const T *y, long n, int op,
                                                 double *part) {
  double acc = 0;
  GL(i, n) {
    const double a = (double)x[i];
    if (op == 0) acc = fmax(acc, fabs(a));
    else if (op == 1) acc += a * (double)y[i];
    else acc += isfinite(a) ? 0.0 : 1.0;
  }
  __shared__ double sh[256];
  sh[threadIdx.x] = acc;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o)
      sh[threadIdx.x] = op == 0 ? fmax(sh[threadIdx.x], sh[threadIdx.x + o])
                                : sh[threadIdx.x] + sh[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}
__global__ void k_reduce2(const double *part, int n, int op, double *out) {
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    double a = 0;
    for (int i = 0; i < n; ++i) a = op == 0 ? fmax(a, part[i]) : a + part[i];
    out[0] = a;
  }
}
// out[w][h] = in[h][w], 32x32 LDS tiles with a padded row
template <typename T>
__global__ void k_transpose(const T *in, long height, long width, T *out) {
  __shared__ T tile[32][33];
  const long x0 = (long)blockIdx.x * 32, y0 = (long)blockIdx.y * 32;
  for (int r = threadIdx.y; r < 32; r += 8) {
    const long y = y0 + r, x = x0 + threadIdx.x;
    if (y < height && x < width) tile[r][threadIdx.x] = in[y * width + x];
  }
  __syncthreads();
  for (int r = threadIdx.y; r < 32; r += 8) {
    const long x = x0 + r, y = y0 + threadIdx.x;
    if (x < width && y < height) out[x * height + y] = tile[threadIdx.x][r];
  }
}

template <typename T>
int reduce_scalar(const T *x, const T *y, long n, int op, double *res) {
  double *ws = (double *)workspace(kWsSolver, sizeof(double) * 1100);
  if (!ws) return VES3D_MEMORY_ERROR;
  const int nb = (int)std::min<long>(1024, std::max<long>(1, (n + 255) / 256));
  k_reduce1<T><<<nb, 256>>>(x, y, n, op, ws);
  V3D_CHECK(launch_status());
  k_reduce2<<<1, 64>>>(ws, nb, op, ws + 1050);
  V3D_CHECK(launch_status());
  return hip_status(hipMemcpy(res, ws + 1050, sizeof(double), hipMemcpyDeviceToHost));
}

std::mutex g_mu;
std::map<std::tuple<int, int, int, long, long, long, long, long, long>, GemmDesc *> g_gemm;

// Device<GPU>::gemm (BLAS column-major, src/DeviceGPU.cc:231-246)
template <typename T>
int blas_gemm(char ta, char tb, int m, int n, int k, T alpha, const T *A, int lda, const T *B,
              int ldb, T beta, T *C, int ldc) {
  if (m < 0 || n < 0 || k < 0) return VES3D_INVALID_PARAMETER;
  if (m == 0 || n == 0) return VES3D_SUCCESS;
  const bool at = (ta == 'T' || ta == 't' || ta == 'C' || ta == 'c');
  const bool bt = (tb == 'T' || tb == 't' || tb == 'C' || tb == 'c');
  const long ars = at ? lda : 1, acs = at ? 1 : lda;
  const long brs = bt ? ldb : 1, bcs = bt ? 1 : ldb;
  GemmDesc h{};
  h.M = m;
  h.N = n;
  h.K = k;
  h.a_rs = ars;
  h.a_cs = acs;
  h.b_rs = brs;
  h.b_cs = bcs;
  h.c_rs = 1;
  h.c_cs = ldc;
  GemmDesc *d = nullptr;
  {
    std::lock_guard<std::mutex> lk(g_mu);
    int dev = 0;
    (void)hipGetDevice(&dev);
    auto key = std::make_tuple(dev, m, n, (long)k, ars, acs, brs, bcs, (long)ldc);
    auto it = g_gemm.find(key);
    if (it != g_gemm.end()) {
      d = it->second;
    } else {
      d = upload_descs(&h, 1);
      if (!d) return VES3D_MEMORY_ERROR;
      g_gemm[key] = d;
    }
  }
  return gemm_grouped<T>(d, 1, 1, m, n, A, B, C, alpha, beta, nullptr, gemm_pick_tile(&h, 1));
}

}  // namespace
}  // namespace v3d

using namespace v3d;

#define V3D_DEVICE_OPS(T, S)                                                                     \
  extern "C" int ves3d_hip_dot_product_##S(const T *u, const T *v, size_t stride, size_t nv,     \
                                           T *out) {                                             \
    if (nv * stride == 0) return VES3D_SUCCESS;                                                  \
    k_dot3<T><<<gb(nv * stride), 256>>>(u, v, stride, nv, out);                                  \
    return launch_status();                                                                      \
  }                                                                                              \
  extern "C" int ves3d_hip_cross_product_##S(const T *u, const T *v, size_t stride, size_t nv,   \
                                             T *out) {                                           \
    if (nv * stride == 0) return VES3D_SUCCESS;                                                  \
    k_cross3<T><<<gb(nv * stride), 256>>>(u, v, stride, nv, out);                                \
    return launch_status();                                                                      \
  }                                                                                              \
  extern "C" int ves3d_hip_sqrt_##S(const T *x, size_t n, T *out) {                              \
    if (n == 0) return VES3D_SUCCESS;                                                            \
    k_sqrt<T><<<gb(n), 256>>>(x, n, out);                                                        \
    return launch_status();                                                                      \
  }                                                                                              \
  extern "C" int ves3d_hip_ax_##S(const T *a, const T *x, size_t stride, size_t nv, T *out) {    \
    if (nv * stride == 0) return VES3D_SUCCESS;                                                  \
    k_ax<T><<<gb(nv * stride), 256>>>(a, x, stride, nv, out);                                    \
    return launch_status();                                                                      \
  }                                                                                              \
  extern "C" int ves3d_hip_xy_##S(const T *x, const T *y, size_t n, T *out) {                    \
    if (n == 0) return VES3D_SUCCESS;                                                            \
    k_xy<T><<<gb(n), 256>>>(x, y, n, out);                                                       \
    return launch_status();                                                                      \
  }                                                                                              \
  extern "C" int ves3d_hip_xyinv_##S(const T *x, const T *y, size_t n, T *out) {                 \
    if (n == 0) return VES3D_SUCCESS;                                                            \
    k_xyinv<T><<<gb(n), 256>>>(x, y, n, out);                                                    \
    return launch_status();                                                                      \
  }                                                                                              \
  extern "C" int ves3d_hip_uyinv_##S(const T *u, const T *y, size_t stride, size_t nv, T *out) { \
    if (nv * stride == 0) return VES3D_SUCCESS;                                                  \
    k_uyinv<T><<<gb(3 * nv * stride), 256>>>(u, y, stride, nv, out);                             \
    return launch_status();                                                                      \
  }                                                                                              \
  extern "C" int ves3d_hip_axpy_##S(T a, const T *x, const T *y, size_t n, T *out) {             \
    if (n == 0) return VES3D_SUCCESS;                                                            \
    k_axpy<T><<<gb(n), 256>>>(a, x, y, n, out);                                                  \
    return launch_status();                                                                      \
  }                                                                                              \
  extern "C" int ves3d_hip_apx_##S(const T *a, const T *x, size_t stride, size_t ns, T *out) {   \
    if (ns * stride == 0) return VES3D_SUCCESS;                                                  \
    k_apx<T><<<gb(ns * stride), 256>>>(a, x, stride, ns, out);                                   \
    return launch_status();                                                                      \
  }                                                                                              \
  extern "C" int ves3d_hip_avpw_##S(const T *a, const T *v, const T *w, size_t stride,           \
                                    size_t nv, T *out) {                                         \
    if (nv * stride == 0) return VES3D_SUCCESS;                                                  \
    k_avpw<T><<<gb(3 * nv * stride), 256>>>(a, v, w, stride, nv, out);                           \
    return launch_status();                                                                      \
  }                                                                                              \
  extern "C" int ves3d_hip_xvpw_##S(const T *x, const T *v, const T *w, size_t stride,           \
                                    size_t nv, T *out) {                                         \
    if (nv * stride == 0) return VES3D_SUCCESS;                                                  \
    k_xvpw<T><<<gb(3 * nv * stride), 256>>>(x, v, w, stride, nv, out);                           \
    return launch_status();                                                                      \
  }                                                                                              \
  extern "C" int ves3d_hip_reduce_##S(const T *x, int x_dim, const T *w, const T *qw,           \
                                      size_t stride, size_t ns, T *out) {                        \
    if (ns == 0) return VES3D_SUCCESS;                                                           \
    if (x_dim < 1 || !w || !qw || !out) return VES3D_INVALID_PARAMETER;                          \
    const int xd = x ? x_dim : 1;                                                                \
    k_reduce<T><<<(unsigned)(ns * xd), 256>>>(x, xd, w, qw, stride, out);                        \
    return launch_status();                                                                      \
  }                                                                                              \
  extern "C" int ves3d_hip_maxabs_##S(const T *x, size_t n, T *result_host) {                    \
    if (!result_host) return VES3D_INVALID_PARAMETER;                                            \
    if (n == 0) {                                                                                \
      *result_host = 0;                                                                          \
      return VES3D_SUCCESS;                                                                      \
    }                                                                                            \
    double r = 0;                                                                                \
    V3D_CHECK(reduce_scalar<T>(x, nullptr, (long)n, 0, &r));                                     \
    *result_host = (T)r;                                                                         \
    return VES3D_SUCCESS;                                                                        \
  }                                                                                              \
  extern "C" int ves3d_hip_algebraic_dot_##S(const T *x, const T *y, size_t n, T *result_host) { \
    if (!result_host) return VES3D_INVALID_PARAMETER;                                            \
    double r = 0;                                                                                \
    if (n) V3D_CHECK(reduce_scalar<T>(x, y, (long)n, 1, &r));                                    \
    *result_host = (T)r;                                                                         \
    return VES3D_SUCCESS;                                                                        \
  }                                                                                              \
  extern "C" int ves3d_hip_isnumeric_##S(const T *x, size_t n, int *result_host) {               \
    if (!result_host) return VES3D_INVALID_PARAMETER;                                            \
    double r = 0;                                                                                \
    if (n) V3D_CHECK(reduce_scalar<T>(x, nullptr, (long)n, 2, &r));                              \
    *result_host = r == 0 ? 1 : 0;                                                               \
    return VES3D_SUCCESS;                                                                        \
  }                                                                                              \
  extern "C" int ves3d_hip_transpose_##S(const T *in, size_t height, size_t width, T *out) {     \
    if (height * width == 0) return VES3D_SUCCESS;                                               \
    if (in == out) return VES3D_INVALID_PARAMETER;                                               \
    dim3 grid(cdiv(width, 32), cdiv(height, 32)), block(32, 8);                                  \
    k_transpose<T><<<grid, block>>>(in, height, width, out);                                     \
    return launch_status();                                                                      \
  }                                                                                              \
  extern "C" int ves3d_hip_gemm_##S(char transA, char transB, int m, int n, int k, T alpha,      \
                                    const T *A, int lda, const T *B, int ldb, T beta, T *C,      \
                                    int ldc) {                                                   \
    return blas_gemm<T>(transA, transB, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc);           \
  }

V3D_DEVICE_OPS(double, f64)
V3D_DEVICE_OPS(float, f32)

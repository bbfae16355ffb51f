// gemm.hip -- grouped strided batched GEMM (see gemm.hpp).
//
// 64x64 output tile per 256-thread workgroup, K staged through LDS in steps
// of 16; each lane owns a 4x4 register tile.  Operand tiles are loaded with
// the lane-to-element mapping chosen from the operand's unit stride so global
// reads stay coalesced whether the operand is row- or column-contiguous.
#include <atomic>
#include <cstdlib>
#include <vector>

#include "gemm.hpp"

namespace v3d {

constexpr int GB_M = 64, GB_N = 64, GB_K = 16, GB_T = 256;

template <typename T>
__global__ __launch_bounds__(GB_T) void k_gemm_grouped(
    const GemmDesc *__restrict__ descs, const T *__restrict__ A,
    const T *__restrict__ B, T *__restrict__ C, T alpha, T beta) {
  __shared__ T As[GB_K][GB_M + 1];
  __shared__ T Bs[GB_K][GB_N + 1];
  const GemmDesc d = descs[blockIdx.y];
  const int tiles_n = (d.N + GB_N - 1) / GB_N;
  const int tm = blockIdx.x / tiles_n, tn = blockIdx.x % tiles_n;
  if (tm * GB_M >= d.M || tn * GB_N >= d.N) return;
  const long b = blockIdx.z;
  const T *Ab = A + d.a_off + b * d.a_bs;
  const T *Bb = B + d.b_off + b * d.b_bs;
  T *Cb = C + d.c_off + b * d.c_bs;
  const int tid = threadIdx.x;
  const int tx = tid % 16, ty = tid / 16;
  const int m0 = tm * GB_M, n0 = tn * GB_N;
  const bool a_kfast = (d.a_cs == 1);
  const bool b_nfast = (d.b_cs == 1);

  T acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = T(0);

  for (int k0 = 0; k0 < d.K; k0 += GB_K) {
#pragma unroll
    for (int q = 0; q < (GB_M * GB_K) / GB_T; ++q) {
      const int e = tid + q * GB_T;
      int m, k;
      if (a_kfast) {
        k = e % GB_K;
        m = e / GB_K;
      } else {
        m = e % GB_M;
        k = e / GB_M;
      }
      const int gm = m0 + m, gk = k0 + k;
      As[k][m] = (gm < d.M && gk < d.K) ? Ab[gm * d.a_rs + gk * d.a_cs] : T(0);
    }
#pragma unroll
    for (int q = 0; q < (GB_N * GB_K) / GB_T; ++q) {
      const int e = tid + q * GB_T;
      int n, k;
      if (b_nfast) {
        n = e % GB_N;
        k = e / GB_N;
      } else {
        k = e % GB_K;
        n = e / GB_K;
      }
      const int gn = n0 + n, gk = k0 + k;
      Bs[k][n] = (gn < d.N && gk < d.K) ? Bb[gk * d.b_rs + gn * d.b_cs] : T(0);
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < GB_K; ++kk) {
      T a[4], bv[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) a[i] = As[kk][ty * 4 + i];
#pragma unroll
      for (int j = 0; j < 4; ++j) bv[j] = Bs[kk][tx * 4 + j];
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] += a[i] * bv[j];
    }
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int gm = m0 + ty * 4 + i;
    if (gm >= d.M) continue;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int gn = n0 + tx * 4 + j;
      if (gn >= d.N) continue;
      T *c = Cb + gm * d.c_rs + gn * d.c_cs;
      *c = (beta == T(0)) ? alpha * acc[i][j] : alpha * acc[i][j] + beta * (*c);
    }
  }
}

// MFMA form of the same tile: the 64x64 output tile is split into 2x2 wave
// sub-tiles of 32x32, each a 2x2 block of 16x16 MFMA accumulators.  For both
// v_mfma_f64_16x16x4f64 and v_mfma_f32_16x16x4f32 lane l supplies A(l%16, l/16)
// and B(l/16, l%16) of a 16x4 / 4x16 step and holds 4 results of column
// l%16; their rows differ: l/16 + 4r (f64) and 4(l/16) + r (f32) for result
// register r (the accumulator layouts composable_kernel's xdlops_gemm.hpp
// records as group_size 1 / 4).  LDS staging is shared with the VALU kernel.
template <typename T>
struct Mfma;
template <>
struct Mfma<double> {
  typedef double v4 __attribute__((ext_vector_type(4)));
  static __device__ __forceinline__ v4 step(double a, double b, v4 c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int lk, int r) { return lk + 4 * r; }
};
template <>
struct Mfma<float> {
  typedef float v4 __attribute__((ext_vector_type(4)));
  static __device__ __forceinline__ v4 step(float a, float b, v4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int lk, int r) { return 4 * lk + r; }
};

template <typename T>
__global__ __launch_bounds__(GB_T) void k_gemm_mfma(const GemmDesc *__restrict__ descs,
                                                    const T *__restrict__ A,
                                                    const T *__restrict__ B, T *__restrict__ C,
                                                    T alpha, T beta) {
  typedef typename Mfma<T>::v4 v4;
  __shared__ T As[GB_K][GB_M + 1];
  __shared__ T Bs[GB_K][GB_N + 1];
  const GemmDesc d = descs[blockIdx.y];
  const int tiles_n = (d.N + GB_N - 1) / GB_N;
  const int tm = blockIdx.x / tiles_n, tn = blockIdx.x % tiles_n;
  if (tm * GB_M >= d.M || tn * GB_N >= d.N) return;
  const long b = blockIdx.z;
  const T *Ab = A + d.a_off + b * d.a_bs;
  const T *Bb = B + d.b_off + b * d.b_bs;
  T *Cb = C + d.c_off + b * d.c_bs;
  const int tid = threadIdx.x;
  const int lane = tid & 63, w = tid >> 6;
  const int li = lane & 15, lk = lane >> 4;
  const int wm = (w >> 1) * 32, wn = (w & 1) * 32;
  const int m0 = tm * GB_M, n0 = tn * GB_N;
  const bool a_kfast = (d.a_cs == 1);
  const bool b_nfast = (d.b_cs == 1);
  v4 acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = v4{T(0), T(0), T(0), T(0)};

  for (int k0 = 0; k0 < d.K; k0 += GB_K) {
#pragma unroll
    for (int q = 0; q < (GB_M * GB_K) / GB_T; ++q) {
      const int e = tid + q * GB_T;
      int m, k;
      if (a_kfast) {
        k = e % GB_K;
        m = e / GB_K;
      } else {
        m = e % GB_M;
        k = e / GB_M;
      }
      const int gm = m0 + m, gk = k0 + k;
      As[k][m] = (gm < d.M && gk < d.K) ? Ab[gm * d.a_rs + gk * d.a_cs] : T(0);
    }
#pragma unroll
    for (int q = 0; q < (GB_N * GB_K) / GB_T; ++q) {
      const int e = tid + q * GB_T;
      int n, k;
      if (b_nfast) {
        n = e % GB_N;
        k = e / GB_N;
      } else {
        k = e % GB_K;
        n = e / GB_K;
      }
      const int gn = n0 + n, gk = k0 + k;
      Bs[k][n] = (gn < d.N && gk < d.K) ? Bb[gk * d.b_rs + gn * d.b_cs] : T(0);
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < GB_K; kk += 4) {
      T a[2], bv[2];
#pragma unroll
      for (int i = 0; i < 2; ++i) a[i] = As[kk + lk][wm + i * 16 + li];
#pragma unroll
      for (int j = 0; j < 2; ++j) bv[j] = Bs[kk + lk][wn + j * 16 + li];
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = Mfma<T>::step(a[i], bv[j], acc[i][j]);
    }
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      const int gn = n0 + wn + j * 16 + li;
      if (gn >= d.N) continue;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int gm = m0 + wm + i * 16 + Mfma<T>::row(lk, r);
        if (gm >= d.M) continue;
        T *c = Cb + gm * d.c_rs + gn * d.c_cs;
        *c = (beta == T(0)) ? alpha * acc[i][j][r] : alpha * acc[i][j][r] + beta * (*c);
      }
    }
}

// Kernel choice: 1 = MFMA (the default), 0 = VALU.  Initialised from
// VES3D_GEMM ("valu" selects the VALU tiles) at load, changed by
// ves3d_hip_set_gemm_kernel.
std::atomic<int> g_gemm_kind{[] {
  const char *e = getenv("VES3D_GEMM");
  return (e && e[0] == 'v') ? 0 : 1;
}()};

bool use_mfma() { return g_gemm_kind.load(std::memory_order_relaxed) == 1; }

template <typename T>
int gemm_grouped(const GemmDesc *d_descs, int ndesc, int batch, int max_m,
                 int max_n, const T *A, const T *B, T *C, T alpha, T beta,
                 hipStream_t st) {
  if (ndesc == 0 || batch == 0 || max_m == 0 || max_n == 0) return VES3D_SUCCESS;
  if (ndesc > 65535 || batch > 65535) return VES3D_INVALID_PARAMETER;
  const long tiles = (long)cdiv(max_m, GB_M) * cdiv(max_n, GB_N);
  if (tiles > 0x7fffffffL) return VES3D_INVALID_PARAMETER;
  dim3 grid((unsigned)tiles, ndesc, batch);
  if (use_mfma())
    k_gemm_mfma<T><<<grid, GB_T, 0, st>>>(d_descs, A, B, C, alpha, beta);
  else
    k_gemm_grouped<T><<<grid, GB_T, 0, st>>>(d_descs, A, B, C, alpha, beta);
  return launch_status();
}

template int gemm_grouped<double>(const GemmDesc *, int, int, int, int, const double *,
                                  const double *, double *, double, double, hipStream_t);
template int gemm_grouped<float>(const GemmDesc *, int, int, int, int, const float *,
                                 const float *, float *, float, float, hipStream_t);

GemmDesc *upload_descs(const GemmDesc *h, int n) {
  GemmDesc *d = nullptr;
  if (hipMalloc(&d, sizeof(GemmDesc) * (n > 0 ? n : 1)) != hipSuccess) return nullptr;
  if (n > 0 && hipMemcpy(d, h, sizeof(GemmDesc) * n, hipMemcpyHostToDevice) != hipSuccess) {
    (void)hipFree(d);
    return nullptr;
  }
  return d;
}

}  // namespace v3d

extern "C" int ves3d_hip_set_gemm_kernel(int kind) {
  if (kind != 0 && kind != 1) return VES3D_INVALID_PARAMETER;
  v3d::g_gemm_kind.store(kind, std::memory_order_relaxed);
  return VES3D_SUCCESS;
}

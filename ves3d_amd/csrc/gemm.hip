// gemm.hip -- grouped strided batched GEMM (see gemm.hpp).
//
// 64x64 output tile per 256-thread workgroup, K staged through LDS in steps
// of 16; each lane owns a 4x4 register tile.  Operand tiles are loaded with
// the lane-to-element mapping chosen from the operand's unit stride so global
// reads stay coalesced whether the operand is row- or column-contiguous.
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

template <typename T>
int gemm_grouped(const GemmDesc *d_descs, int ndesc, int batch, int max_m,
                 int max_n, const T *A, const T *B, T *C, T alpha, T beta,
                 hipStream_t st) {
  if (ndesc == 0 || batch == 0 || max_m == 0 || max_n == 0) return VES3D_SUCCESS;
  if (ndesc > 65535 || batch > 65535) return VES3D_INVALID_PARAMETER;
  const long tiles = (long)cdiv(max_m, GB_M) * cdiv(max_n, GB_N);
  if (tiles > 0x7fffffffL) return VES3D_INVALID_PARAMETER;
  dim3 grid((unsigned)tiles, ndesc, batch);
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

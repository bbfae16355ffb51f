// gemm.hip -- grouped strided batched GEMM (see gemm.hpp).
//
// Two kernels behind one entry.  The MFMA kernel (the default) takes its
// output tile shape from the descriptor list's plan (64 or 128 rows by 16-64
// columns; gemm_pick_tile).  The VALU kernel has a 64x64 tile in which each
// lane owns a 4x4 register block.  Both stage K through LDS in slabs of 16,
// with the lane-to-element mapping chosen from the operand's unit stride so
// global reads stay coalesced whether the operand is row- or
// column-contiguous.
#include <atomic>
#include <cstdint>
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

// MFMA operand layout.  For both v_mfma_f64_16x16x4f64 and
// v_mfma_f32_16x16x4f32 lane l supplies A(l%16, l/16) and B(l/16, l%16) of a
// 16x4 / 4x16 step and holds 4 results of column l%16; their rows differ:
// l/16 + 4r (f64) and 4(l/16) + r (f32) for result register r (the
// accumulator layouts composable_kernel's xdlops_gemm.hpp records as
// group_size 1 / 4).
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

// MFMA output tile of BM = 64 NTM rows by BN = 16 NTN columns: wave w owns
// rows [16 NTM w, 16 NTM (w + 1)) and all BN columns, as NTM x NTN MFMA
// accumulators.  Operands are staged through LDS in 16-deep K slabs (the
// same coalesced lane mapping as the VALU kernel), the next slab's loads in
// flight while the current slab's MFMAs run.  The K loop runs the MFMA steps
// only up to the descriptor's K (rounded up to 4), and a wave whose rows all
// lie past M skips its steps.
//
// AV: the A operand is read as 16-byte vectors along its contiguous direction
// (k for a row-contiguous A, m for a column-contiguous one).  The host sets it
// only when every descriptor keeps those vectors aligned and whole (offsets,
// strides and the contiguous extent multiples of the vector length, base
// pointer 16-byte aligned): the Fourier stages' operands.
template <typename T, int NTM, int NTN, bool AV>
__global__ __launch_bounds__(GB_T) void k_gemm_mfma(const GemmDesc *__restrict__ descs,
                                                    const T *__restrict__ A,
                                                    const T *__restrict__ B, T *__restrict__ C,
                                                    T alpha, T beta) {
  typedef typename Mfma<T>::v4 v4;
  constexpr int BM = 64 * NTM, BN = 16 * NTN;
  constexpr int VEC = AV ? 16 / (int)sizeof(T) : 1;
  typedef T vec_t __attribute__((ext_vector_type(VEC)));
  static_assert((BM * GB_K) % GB_T == 0 && (BN * GB_K) % GB_T == 0, "tile loads");
  __shared__ T As[GB_K][BM + 1];
  __shared__ T Bs[GB_K][BN + 1];
  const GemmDesc d = descs[blockIdx.y];
  const int tiles_n = (d.N + BN - 1) / BN;
  const int tm = blockIdx.x / tiles_n, tn = blockIdx.x % tiles_n;
  if (tm * BM >= d.M || tn * BN >= d.N) return;
  const long b = blockIdx.z;
  const T *Ab = A + d.a_off + b * d.a_bs;
  const T *Bb = B + d.b_off + b * d.b_bs;
  T *Cb = C + d.c_off + b * d.c_bs;
  const int tid = threadIdx.x;
  const int lane = tid & 63, w = tid >> 6;
  const int li = lane & 15, lk = lane >> 4;
  const int wm = w * 16 * NTM;
  const int m0 = tm * BM, n0 = tn * BN;
  const bool a_kfast = (d.a_cs == 1);
  const bool b_nfast = (d.b_cs == 1);
  const bool wave_live = m0 + wm < d.M;
  v4 acc[NTM][NTN];
#pragma unroll
  for (int i = 0; i < NTM; ++i)
#pragma unroll
    for (int j = 0; j < NTN; ++j) acc[i][j] = v4{T(0), T(0), T(0), T(0)};

  // The next slab's operands are loaded into registers while the current
  // slab's MFMAs run.
  constexpr int QA = (BM * GB_K) / GB_T, QB = (BN * GB_K) / GB_T;
  T ra[QA], rb[QB];
  auto a_idx = [&](int q, int &m, int &k) {
    const int e = tid + q * GB_T;
    if (a_kfast) {
      k = e % GB_K;
      m = e / GB_K;
    } else {
      m = e % BM;
      k = e / BM;
    }
  };
  auto b_idx = [&](int q, int &n, int &k) {
    const int e = tid + q * GB_T;
    if (b_nfast) {
      n = e % BN;
      k = e / BN;
    } else {
      k = e % GB_K;
      n = e / GB_K;
    }
  };
  // vector q of the A slab: first element (m, k), vector along k (row-
  // contiguous A) or along m
  auto a_vidx = [&](int q, int &m, int &k) {
    const int e = tid + q * GB_T;
    if (a_kfast) {
      k = (e % (GB_K / VEC)) * VEC;
      m = e / (GB_K / VEC);
    } else {
      m = (e % (BM / VEC)) * VEC;
      k = e / (BM / VEC);
    }
  };
  auto gload = [&](int k0) {
    if constexpr (AV) {
#pragma unroll
      for (int q = 0; q < QA / VEC; ++q) {
        int m, k;
        a_vidx(q, m, k);
        const int gm = m0 + m, gk = k0 + k;
        vec_t v = (gm < d.M && gk < d.K)
                      ? *(const vec_t *)(Ab + gm * d.a_rs + gk * d.a_cs)
                      : vec_t(T(0));
#pragma unroll
        for (int u = 0; u < VEC; ++u) ra[q * VEC + u] = v[u];
      }
    } else {
#pragma unroll
      for (int q = 0; q < QA; ++q) {
        int m, k;
        a_idx(q, m, k);
        const int gm = m0 + m, gk = k0 + k;
        ra[q] = (gm < d.M && gk < d.K) ? Ab[gm * d.a_rs + gk * d.a_cs] : T(0);
      }
    }
#pragma unroll
    for (int q = 0; q < QB; ++q) {
      int n, k;
      b_idx(q, n, k);
      const int gn = n0 + n, gk = k0 + k;
      rb[q] = (gn < d.N && gk < d.K) ? Bb[gk * d.b_rs + gn * d.b_cs] : T(0);
    }
  };
  gload(0);
  for (int k0 = 0; k0 < d.K; k0 += GB_K) {
    if constexpr (AV) {
#pragma unroll
      for (int q = 0; q < QA / VEC; ++q) {
        int m, k;
        a_vidx(q, m, k);
#pragma unroll
        for (int u = 0; u < VEC; ++u) {
          if (a_kfast) As[k + u][m] = ra[q * VEC + u];
          else As[k][m + u] = ra[q * VEC + u];
        }
      }
    } else {
#pragma unroll
      for (int q = 0; q < QA; ++q) {
        int m, k;
        a_idx(q, m, k);
        As[k][m] = ra[q];
      }
    }
#pragma unroll
    for (int q = 0; q < QB; ++q) {
      int n, k;
      b_idx(q, n, k);
      Bs[k][n] = rb[q];
    }
    __syncthreads();
    if (k0 + GB_K < d.K) gload(k0 + GB_K);
    const int kmax = d.K - k0;
    if (wave_live) {
#pragma unroll
      for (int kk = 0; kk < GB_K; kk += 4) {
        if (kk < kmax) {
          T a[NTM], bv[NTN];
#pragma unroll
          for (int i = 0; i < NTM; ++i) a[i] = As[kk + lk][wm + i * 16 + li];
#pragma unroll
          for (int j = 0; j < NTN; ++j) bv[j] = Bs[kk + lk][j * 16 + li];
#pragma unroll
          for (int i = 0; i < NTM; ++i)
#pragma unroll
            for (int j = 0; j < NTN; ++j) acc[i][j] = Mfma<T>::step(a[i], bv[j], acc[i][j]);
        }
      }
    }
    __syncthreads();
  }
  if (d.c_rs == 1 && d.c_cs != 1) {
    // Column-contiguous C (the Fourier stages write their output transposed):
    // the accumulators hold 16 columns x 4 rows per store instruction, 16
    // short runs.  Each 16-column slice of the tile goes through LDS instead
    // and leaves as whole columns, consecutive lanes on consecutive rows.
#pragma unroll
    for (int j = 0; j < NTN; ++j) {
      __syncthreads();
      if (wave_live) {
#pragma unroll
        for (int i = 0; i < NTM; ++i)
#pragma unroll
          for (int r = 0; r < 4; ++r) As[li][wm + i * 16 + Mfma<T>::row(lk, r)] = acc[i][j][r];
      }
      __syncthreads();
#pragma unroll
      for (int q = 0; q < (16 * BM) / GB_T; ++q) {
        const int e = tid + q * GB_T;
        const int rr = e % BM, cc = e / BM;
        const int gm = m0 + rr, gn = n0 + j * 16 + cc;
        if (gm < d.M && gn < d.N) {
          T *c = Cb + gm + gn * d.c_cs;
          *c = (beta == T(0)) ? alpha * As[cc][rr] : alpha * As[cc][rr] + beta * (*c);
        }
      }
    }
    return;
  }
  if (!wave_live) return;
#pragma unroll
  for (int i = 0; i < NTM; ++i)
#pragma unroll
    for (int j = 0; j < NTN; ++j) {
      const int gn = n0 + j * 16 + li;
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

// MFMA tile shapes (NTM, NTN); index 0 is the 64x64 tile.  A tile choice is
// the shape index plus flags saying the list's A operand allows 16-byte
// vector loads for 8-byte and for 4-byte elements.
constexpr int kTiles[][2] = {{1, 4}, {1, 1}, {2, 1}, {4, 1}, {1, 2},
                             {2, 2}, {1, 3}, {2, 3}, {2, 4}};
constexpr int kNumTiles = sizeof(kTiles) / sizeof(kTiles[0]);
constexpr int kTileShapeMask = 15, kTileVec8 = 16, kTileVec4 = 32;

int gemm_pick_tile(const GemmDesc *h, int n) {
  // Cost of a list in multiply-add equivalents: the padded multiply-adds of
  // whole tiles (a partly filled tile holds its block's residency even where
  // its waves idle; K rounded up to 4), the A operand read once per column
  // tile and C written once, at 39 multiply-adds per element (the MFMA rate
  // over the HBM rate, the same in FP64 and FP32), and a fixed 16384 per
  // tile for the prologue and the barriers.  The 256-row shape is left to
  // forced selection: its 35 KB of FP64 LDS per block measured slower than
  // the 64- and 128-row shapes on every SH / rotation plan.
  int best = 0;
  double best_cost = -1;
  for (int t = 0; t < kNumTiles; ++t) {
    if (kTiles[t][0] > 2) continue;
    const int BM = 64 * kTiles[t][0], BN = 16 * kTiles[t][1];
    double cost = 0;
    for (int i = 0; i < n; ++i) {
      const GemmDesc &d = h[i];
      if (d.M <= 0 || d.N <= 0 || d.K <= 0) continue;
      const long tiles_m = cdiv(d.M, BM), tiles_n = cdiv(d.N, BN);
      cost += (double)tiles_m * BM * tiles_n * BN * (4 * cdiv(d.K, 4)) +
              39.0 * ((double)d.M * d.K * tiles_n + (double)d.M * d.N) +
              16384.0 * tiles_m * tiles_n;
    }
    if (best_cost < 0 || cost < best_cost) {
      best_cost = cost;
      best = t;
    }
  }
  // whole aligned 16-byte A vectors for 8-byte (2 elements) / 4-byte (4) types
  auto vec_ok = [&](long v) {
    for (int i = 0; i < n; ++i) {
      const GemmDesc &d = h[i];
      if (d.a_off % v || d.a_bs % v) return false;
      if (d.a_cs == 1) {
        if (d.a_rs % v || d.K % v) return false;
      } else if (d.a_rs == 1) {
        if (d.a_cs % v || d.M % v) return false;
      } else {
        return false;
      }
    }
    return n > 0;
  };
  return best | (vec_ok(2) ? kTileVec8 : 0) | (vec_ok(4) ? kTileVec4 : 0);
}

// Kernel choice: 1 = MFMA (the default), 0 = VALU.  Initialised from
// VES3D_GEMM ("valu" selects the VALU tiles) at load, changed by
// ves3d_hip_set_gemm_kernel.
std::atomic<int> g_gemm_kind{[] {
  const char *e = getenv("VES3D_GEMM");
  return (e && e[0] == 'v') ? 0 : 1;
}()};

bool use_mfma() { return g_gemm_kind.load(std::memory_order_relaxed) == 1; }

// Forced MFMA tile (-1: the caller's choice), set by ves3d_hip_set_gemm_tile.
std::atomic<int> g_gemm_tile{-1};

template <typename T, int NTM, int NTN>
int launch_mfma(bool av, const GemmDesc *d_descs, int ndesc, int batch, int max_m, int max_n,
                const T *A, const T *B, T *C, T alpha, T beta, hipStream_t st) {
  const long tiles = (long)cdiv(max_m, 64 * NTM) * cdiv(max_n, 16 * NTN);
  if (tiles > 0x7fffffffL) return VES3D_INVALID_PARAMETER;
  dim3 grid((unsigned)tiles, ndesc, batch);
  if (av)
    k_gemm_mfma<T, NTM, NTN, true><<<grid, GB_T, 0, st>>>(d_descs, A, B, C, alpha, beta);
  else
    k_gemm_mfma<T, NTM, NTN, false><<<grid, GB_T, 0, st>>>(d_descs, A, B, C, alpha, beta);
  return launch_status();
}

template <typename T>
int gemm_grouped(const GemmDesc *d_descs, int ndesc, int batch, int max_m,
                 int max_n, const T *A, const T *B, T *C, T alpha, T beta,
                 hipStream_t st, int tile) {
  if (ndesc == 0 || batch == 0 || max_m == 0 || max_n == 0) return VES3D_SUCCESS;
  if (ndesc > 65535 || batch > 65535) return VES3D_INVALID_PARAMETER;
  if (tile < -1 || (tile >= 0 && (tile & kTileShapeMask) >= kNumTiles))
    return VES3D_INVALID_PARAMETER;
  const bool av = tile >= 0 && (tile & (sizeof(T) == 8 ? kTileVec8 : kTileVec4)) &&
                  ((uintptr_t)A & 15) == 0;
  int shape = tile < 0 ? 0 : (tile & kTileShapeMask);
  const int forced = g_gemm_tile.load(std::memory_order_relaxed);
  if (forced >= 0) shape = forced;
  if (!use_mfma()) {
    const long tiles = (long)cdiv(max_m, GB_M) * cdiv(max_n, GB_N);
    if (tiles > 0x7fffffffL) return VES3D_INVALID_PARAMETER;
    dim3 grid((unsigned)tiles, ndesc, batch);
    k_gemm_grouped<T><<<grid, GB_T, 0, st>>>(d_descs, A, B, C, alpha, beta);
    return launch_status();
  }
#define V3D_TILE(I, NTM, NTN)                                                                  \
  case I:                                                                                      \
    return launch_mfma<T, NTM, NTN>(av, d_descs, ndesc, batch, max_m, max_n, A, B, C, alpha, \
                                    beta, st);
  switch (shape) {
    V3D_TILE(0, 1, 4)
    V3D_TILE(1, 1, 1)
    V3D_TILE(2, 2, 1)
    V3D_TILE(3, 4, 1)
    V3D_TILE(4, 1, 2)
    V3D_TILE(5, 2, 2)
    V3D_TILE(6, 1, 3)
    V3D_TILE(7, 2, 3)
    V3D_TILE(8, 2, 4)
  }
#undef V3D_TILE
  return VES3D_INVALID_PARAMETER;
}

template int gemm_grouped<double>(const GemmDesc *, int, int, int, int, const double *,
                                  const double *, double *, double, double, hipStream_t, int);
template int gemm_grouped<float>(const GemmDesc *, int, int, int, int, const float *,
                                 const float *, float *, float, float, hipStream_t, int);

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

extern "C" int ves3d_hip_set_gemm_tile(int tile) {
  if (tile < -1 || tile >= v3d::kNumTiles) return VES3D_INVALID_PARAMETER;
  v3d::g_gemm_tile.store(tile, std::memory_order_relaxed);
  return VES3D_SUCCESS;
}

extern "C" int ves3d_hip_set_gemm_kernel(int kind) {
  if (kind != 0 && kind != 1) return VES3D_INVALID_PARAMETER;
  v3d::g_gemm_kind.store(kind, std::memory_order_relaxed);
  return VES3D_SUCCESS;
}

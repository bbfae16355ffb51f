// learned_pc.hip -- forward / backward / Adam kernels of the learned
// SH-space preconditioner (see learned_pc.hpp) and its hooks into the
// Galerkin solver vector.
#include <algorithm>
#include <cmath>
#include <random>
#include <vector>

#include "learned_pc.hpp"
#include "runtime.hpp"

namespace v3d {

namespace {
inline unsigned gb(long n) {
  long b = (n + 255) / 256;
  if (b > 65536) b = 65536;
  return (unsigned)(b < 1 ? 1 : b);
}
#define GL(i, n) \
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < (n); i += (long)gridDim.x * blockDim.x)

constexpr float kC0 = 0.7978845608028654f;  // sqrt(2/pi)
constexpr float kC1 = 0.044715f;

__device__ __forceinline__ float gelu_f(float x) {
  return 0.5f * x * (1.0f + tanhf(kC0 * (x + kC1 * x * x * x)));
}
__device__ __forceinline__ float gelu_g(float x) {
  const float t = tanhf(kC0 * (x + kC1 * x * x * x));
  return 0.5f * (1.0f + t) + 0.5f * x * (1.0f - t * t) * kC0 * (1.0f + 3.0f * kC1 * x * x);
}

// splitmix64 -> uniform in [-1, 1)
__device__ __forceinline__ float hash_uniform(unsigned long long s) {
  s += 0x9E3779B97F4A7C15ULL;
  s = (s ^ (s >> 30)) * 0xBF58476D1CE4E5B9ULL;
  s = (s ^ (s >> 27)) * 0x94D049BB133111EBULL;
  s ^= s >> 31;
  return (float)((double)(s >> 11) * (1.0 / 9007199254740992.0) * 2.0 - 1.0);
}

__global__ void k_init_uniform(float *p, long n, float a, unsigned long long seed) {
  GL(i, n) p[i] = a * hash_uniform(seed * 0x100000001B3ULL + (unsigned long long)i);
}
__global__ void k_bias_gelu(float *A, const float *__restrict__ b, long B, int w, float *H) {
  GL(i, B * w) {
    const float a = A[i] + b[i % w];
    A[i] = a;
    H[i] = gelu_f(a);
  }
}
__global__ void k_add_bias(float *Y, const float *__restrict__ b, long B, int w) {
  GL(i, B * w) Y[i] += b[i % w];
}
// dY = (Y - T) / n, per-block partial sums of (Y - T)^2 in FP64
__global__ __launch_bounds__(256) void k_diff_loss(const float *__restrict__ Y,
                                                   const float *__restrict__ T, long n, float invn,
                                                   float *dY, double *part) {
  double acc = 0;
  GL(i, n) {
    const float d = Y[i] - T[i];
    dY[i] = d * invn;
    acc += (double)d * d;
  }
  __shared__ double sh[4];
  acc = wave_sum(acc);
  if (threadIdx.x % 64 == 0) sh[threadIdx.x / 64] = acc;
  __syncthreads();
  if (threadIdx.x == 0) part[blockIdx.x] = sh[0] + sh[1] + sh[2] + sh[3];
}
__global__ void k_sum_parts(const double *part, int n, double scale, double *out) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    double s = 0;
    for (int i = 0; i < n; ++i) s += part[i];
    out[0] = s * scale;
  }
}
__global__ void k_gelu_back(float *dH, const float *__restrict__ A, long n) {
  GL(i, n) dH[i] *= gelu_g(A[i]);
}
__global__ void k_colsum(const float *__restrict__ D, long B, int w, float *g) {
  GL(c, (long)w) {
    float s = 0;
    for (long r = 0; r < B; ++r) s += D[r * w + c];
    g[c] = s;
  }
}
// ---- BF16 activations on MFMA
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// y = bf16(x), round to nearest even
__global__ void k_to_bf16(const float *__restrict__ x, long n, __bf16 *y) {
  GL(i, n) y[i] = (__bf16)x[i];
}
// A += b (FP32, kept for the backward pass); H = bf16(gelu(A))
__global__ void k_bias_gelu_bf(float *A, const float *__restrict__ b, long B, int w, __bf16 *H) {
  GL(i, B * w) {
    const float a = A[i] + b[i % w];
    A[i] = a;
    H[i] = (__bf16)gelu_f(a);
  }
}
// dA *= gelu'(A) (FP32, for the bias gradient); dAh = bf16(dA)
__global__ void k_gelu_back_bf(float *dA, const float *__restrict__ A, long n, __bf16 *dAh) {
  GL(i, n) {
    const float g = dA[i] * gelu_g(A[i]);
    dA[i] = g;
    dAh[i] = (__bf16)g;
  }
}

// C[M][N] (FP32, row stride ldc) = sum_k A(m, k) B(k, n) with BF16 operands
// on v_mfma_f32_16x16x32_bf16 and FP32 accumulation.
//   A(m, k) = A[m lda + k] (AK: k contiguous) or A[k lda + m];
//   B(k, n) = B[n ldb + k] (BK: k contiguous) or B[k ldb + n].
// A workgroup of 4 waves computes a 64 x 64 tile, each wave 32 x 32 as 2 x 2
// MFMA tiles, over K slabs of 32 staged through LDS with k contiguous, so
// that a lane's 8 operands are one 16-byte LDS read: lane l supplies row
// (l % 16) of A and column (l % 16) of B at the slab's k = 8 (l / 16) .. +7,
// the same k for both operands (the sum does not depend on which k the
// instruction assigns to which register, only on A and B agreeing), and
// holds results of rows 4 (l / 16) + r, column l % 16 (r = 0..3).  Global
// loads are 16-byte vectors along the contiguous direction; the host
// requires that extent and the leading dimension to be multiples of 8 and
// both bases 16-byte aligned, so a vector is either whole in range or whole
// out of it (then zero).
constexpr int kBT = 64, kBK = 32, kBPad = 8;
template <bool AK, bool BK>
__global__ __launch_bounds__(256) void k_gemm_bf16(int M, int N, int K,
                                                   const __bf16 *__restrict__ A, long lda,
                                                   const __bf16 *__restrict__ B, long ldb,
                                                   float *__restrict__ C, long ldc) {
  __shared__ __attribute__((aligned(16))) __bf16 As[kBT][kBK + kBPad];
  __shared__ __attribute__((aligned(16))) __bf16 Bs[kBT][kBK + kBPad];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const long m0 = (long)blockIdx.y * kBT, n0 = (long)blockIdx.x * kBT;
  const int wm = (w >> 1) * 32, wn = (w & 1) * 32;
  const int li = lane & 15, lg = lane >> 4;
  f32x4 acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
  const bf16x8 zero8 = {};
  for (int k0 = 0; k0 < K; k0 += kBK) {
    // one 8-vector of each operand per thread: 64 x 32 = 256 vectors
    if constexpr (AK) {
      const int m = tid >> 2, kk = (tid & 3) * 8;
      const long gm = m0 + m, gk = k0 + kk;
      const bf16x8 v = (gm < M && gk < K) ? *(const bf16x8 *)(A + gm * lda + gk) : zero8;
      *(bf16x8 *)&As[m][kk] = v;
    } else {
      const int kk = tid >> 3, mm = (tid & 7) * 8;
      const long gm = m0 + mm, gk = k0 + kk;
      const bf16x8 v = (gm < M && gk < K) ? *(const bf16x8 *)(A + gk * lda + gm) : zero8;
#pragma unroll
      for (int u = 0; u < 8; ++u) As[mm + u][kk] = v[u];
    }
    if constexpr (BK) {
      const int n = tid >> 2, kk = (tid & 3) * 8;
      const long gn = n0 + n, gk = k0 + kk;
      const bf16x8 v = (gn < N && gk < K) ? *(const bf16x8 *)(B + gn * ldb + gk) : zero8;
      *(bf16x8 *)&Bs[n][kk] = v;
    } else {
      const int kk = tid >> 3, nn = (tid & 7) * 8;
      const long gn = n0 + nn, gk = k0 + kk;
      const bf16x8 v = (gn < N && gk < K) ? *(const bf16x8 *)(B + gk * ldb + gn) : zero8;
#pragma unroll
      for (int u = 0; u < 8; ++u) Bs[nn + u][kk] = v[u];
    }
    __syncthreads();
    bf16x8 a[2], b[2];
#pragma unroll
    for (int i = 0; i < 2; ++i) a[i] = *(const bf16x8 *)&As[wm + 16 * i + li][8 * lg];
#pragma unroll
    for (int j = 0; j < 2; ++j) b[j] = *(const bf16x8 *)&Bs[wn + 16 * j + li][8 * lg];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j)
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const long gm = m0 + wm + 16 * i + 4 * lg + r, gn = n0 + wn + 16 * j + li;
        if (gm < M && gn < N) C[gm * ldc + gn] = acc[i][j][r];
      }
}

template <bool AK, bool BK>
int gemm_bf16(int M, int N, int K, const __bf16 *A, long lda, const __bf16 *B, long ldb, float *C,
              long ldc, hipStream_t st) {
  if (M < 1 || N < 1 || K < 1) return VES3D_INVALID_PARAMETER;
  if ((AK ? K : M) % 8 || lda % 8 || (BK ? K : N) % 8 || ldb % 8) return VES3D_INVALID_PARAMETER;
  if (((uintptr_t)A | (uintptr_t)B) & 15) return VES3D_INVALID_PARAMETER;
  const long gy = (M + kBT - 1) / kBT;
  if (gy > 65535) return VES3D_INVALID_PARAMETER;
  dim3 grid((unsigned)((N + kBT - 1) / kBT), (unsigned)gy);
  k_gemm_bf16<AK, BK><<<grid, 256, 0, st>>>(M, N, K, A, lda, B, ldb, C, ldc);
  return launch_status();
}

// Adam (Kingma & Ba 2015, Algorithm 1)
// out = -in
__global__ void k_negate(const float *__restrict__ in, long n, float *out) {
  GL(i, n) out[i] = -in[i];
}
// Y = (Y - Yneg) / 2: the odd part g(z) - g(-z) of the network, so that the
// preconditioner satisfies P(-r) = -P(r) (with the row scaling, P(a r) =
// a P(r) for every real a)
__global__ void k_odd_part(float *Y, const float *__restrict__ Yneg, long n) {
  GL(i, n) Y[i] = 0.5f * (Y[i] - Yneg[i]);
}
// minibatch rows Zb[i] = sgn Zs[perm[i]], Tb[i] = sgn Ts[perm[i]], with
// sgn = -1 on odd batch rows: the samples enter with both signs, as the
// target map is odd
__global__ void k_gather_batch(const float *__restrict__ Zs, const float *__restrict__ Ts,
                               const int *__restrict__ perm, long B, int nin, float *Zb, float *Tb) {
  GL(i, B * nin) {
    const long r = i / nin;
    const long src = (long)perm[r] * nin + (i % nin);
    const float sg = (r & 1) ? -1.0f : 1.0f;
    Zb[i] = sg * Zs[src];
    Tb[i] = sg * Ts[src];
  }
}
// partial sums of (Y - T)^2 and T^2 per block into part[2 * block + {0, 1}]
__global__ __launch_bounds__(256) void k_sq_sums(const float *__restrict__ Y,
                                                 const float *__restrict__ T, long n,
                                                 double *part) {
  double a = 0, b = 0;
  GL(i, n) {
    const double d = (double)Y[i] - (double)T[i];
    a += d * d;
    b += (double)T[i] * (double)T[i];
  }
  __shared__ double sh[2][4];
  a = wave_sum(a);
  b = wave_sum(b);
  if (threadIdx.x % 64 == 0) {
    sh[0][threadIdx.x / 64] = a;
    sh[1][threadIdx.x / 64] = b;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = sh[0][0] + sh[0][1] + sh[0][2] + sh[0][3];
    part[2 * blockIdx.x + 1] = sh[1][0] + sh[1][1] + sh[1][2] + sh[1][3];
  }
}
// acc[0..1] += sums of the nb partial pairs (one thread, fixed order)
__global__ void k_acc_pairs(const double *part, int nb, double *acc) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  double a = 0, b = 0;
  for (int k = 0; k < nb; ++k) {
    a += part[2 * k];
    b += part[2 * k + 1];
  }
  acc[0] += a;
  acc[1] += b;
}
// Adam (Kingma & Ba 2015, Algorithm 1) with decoupled weight decay wd
// (Loshchilov & Hutter 2019): P -= lr (m_hat / (sqrt(v_hat) + eps) + wd P)
__global__ void k_adam(float *P, const float *__restrict__ G, float *M, float *V, long n, float lr,
                       float c1, float c2, float wd) {
  GL(i, n) {
    const float g = G[i];
    const float m = 0.9f * M[i] + 0.1f * g;
    const float v = 0.999f * V[i] + 0.001f * g * g;
    M[i] = m;
    V[i] = v;
    P[i] -= lr * ((m / c1) / (sqrtf(v / c2) + 1e-8f) + wd * P[i]);
  }
}

// Row element j of vesicle v inside the order-major Galerkin vector.
__device__ __forceinline__ long row_index(long v, int j, long nv, int nc, const int2 *cinfo) {
  long base, f, nf;
  int c;
  if (j < 3 * nc) {
    const int d = j / nc;
    c = j % nc;
    f = 3 * v + d;
    nf = 3 * nv;
    base = 0;
  } else {
    c = j - 3 * nc;
    f = v;
    nf = nv;
    base = 3 * nv * (long)nc;
  }
  const int2 bl = cinfo[c];
  return base + nf * bl.x + f * bl.y + (c - bl.x);
}

// scale[v] = |row_v(y)|, one workgroup per vesicle
__global__ __launch_bounds__(256) void k_row_norms(const double *__restrict__ y, long nv, int nc,
                                                   const int2 *__restrict__ cinfo, double *scale) {
  const long v = blockIdx.x;
  double acc = 0;
  for (int j = threadIdx.x; j < 4 * nc; j += 256) {
    const double a = y[row_index(v, j, nv, nc, cinfo)];
    acc += a * a;
  }
  __shared__ double sh[4];
  acc = wave_sum(acc);
  if (threadIdx.x % 64 == 0) sh[threadIdx.x / 64] = acc;
  __syncthreads();
  if (threadIdx.x == 0) scale[v] = sqrt(sh[0] + sh[1] + sh[2] + sh[3]) + 1e-300;
}
// Frames F_v (row-major 3x3 per vesicle, columns the vesicle's canonical
// axes; nullptr: identity) rotate the three velocity blocks of a vesicle's
// row into its canonical frame, z_d = sum_e F[e][d] x_e, and back,
// x_e = sum_d F[e][d] z_d.  The tension block is a scalar and not rotated.
__device__ __forceinline__ double to_canonical(const double *__restrict__ y, long v, int j, long nv,
                                               int nc, const int2 *__restrict__ cinfo,
                                               const double *__restrict__ frames) {
  if (j >= 3 * nc || frames == nullptr) return y[row_index(v, j, nv, nc, cinfo)];
  const int d = j / nc, c = j % nc;
  const double *F = frames + 9 * v;
  double z = 0;
#pragma unroll
  for (int e = 0; e < 3; ++e) z += F[3 * e + d] * y[row_index(v, e * nc + c, nv, nc, cinfo)];
  return z;
}

__global__ void k_gather_rows(const double *__restrict__ y, long nv, int nc,
                              const int2 *__restrict__ cinfo, const double *__restrict__ scale,
                              const double *__restrict__ frames, float *Z) {
  const int nin = 4 * nc;
  GL(i, nv * nin) {
    const long v = i / nin;
    const int j = (int)(i % nin);
    Z[i] = (float)(to_canonical(y, v, j, nv, nc, cinfo, frames) / scale[v]);
  }
}
// The correction skips the tension coefficients of degree > tcut: the
// surface divergence is low-pass filtered at that degree (src/Surface.cc:372),
// so those coefficients span the null space of the implicit operator, and
// the solution the diagonal-preconditioned solve reaches has them at zero.
__global__ void k_scatter_add(const float *__restrict__ Y, long nv, int nc, int p, int tcut,
                              const int2 *__restrict__ cinfo, const double *__restrict__ scale,
                              const double *__restrict__ frames, double *y) {
  const int nin = 4 * nc;
  GL(i, nv * nin) {
    const long v = i / nin;
    const int j = (int)(i % nin);
    double add;
    if (j >= 3 * nc) {
      const int c = j - 3 * nc;
      const int2 bl = cinfo[c];
      const int degree = (p + 1 - bl.y) + (c - bl.x);  // order m = p + 1 - block length
      if (degree > tcut) continue;
      add = (double)Y[i];
    } else if (frames == nullptr) {
      add = (double)Y[i];
    } else {
      const int e = j / nc, c = j % nc;
      const double *F = frames + 9 * v;
      const float *Yv = Y + v * nin;
      add = F[3 * e] * (double)Yv[c] + F[3 * e + 1] * (double)Yv[nc + c] +
            F[3 * e + 2] * (double)Yv[2 * nc + c];
    }
    y[row_index(v, j, nv, nc, cinfo)] += scale[v] * add;
  }
}
// samples in the canonical frames: Z = F^T Dr_v / s, T = F^T (e_v - Dr_v) / s,
// written at rows slot0 + v (mod cap)
__global__ void k_record(const double *__restrict__ Dr, const double *__restrict__ e, long nv,
                         int nc, const int2 *__restrict__ cinfo, const double *__restrict__ scale,
                         const double *__restrict__ frames, long slot0, long cap, float *Zs,
                         float *Ts) {
  const int nin = 4 * nc;
  GL(i, nv * nin) {
    const long v = i / nin;
    const int j = (int)(i % nin);
    const long row = (slot0 + v) % cap;
    const double s = scale[v];
    const double z = to_canonical(Dr, v, j, nv, nc, cinfo, frames);
    const double t = to_canonical(e, v, j, nv, nc, cinfo, frames);
    Zs[row * nin + j] = (float)(z / s);
    Ts[row * nin + j] = (float)((t - z) / s);
  }
}
}  // namespace

LearnedPC::~LearnedPC() {
  for (float *p : {P_, G_, M_, V_, Zs_, Ts_, A1_, H1_, A2_, H2_, Y_, dY_, dA2_, dA1_, Zb_, Tb_,
                   Zneg_, Yneg_})
    if (p) (void)hipFree(p);
  if (scale_) (void)hipFree(scale_);
  if (lossbuf_) (void)hipFree(lossbuf_);
  if (evalbuf_) (void)hipFree(evalbuf_);
  if (perm_) (void)hipFree(perm_);
  if (vidx_) (void)hipFree(vidx_);
  if (Pbest_) (void)hipFree(Pbest_);
  for (__bf16 *p : {Wh_, Zh_, H1h_, H2h_, dYh_, dA2h_, dA1h_})
    if (p) (void)hipFree(p);
  for (auto &kv : descs_) (void)hipFree(kv.second);
}

int LearnedPC::set_activations(int mode) {
  if (mode != 0 && mode != 1) return VES3D_INVALID_PARAMETER;
  if (mode == 1 && (nin_ % 8 || width_ % 8)) return VES3D_INVALID_PARAMETER;
  if (mode == 1 && !Wh_) {
    const long nw = (long)width_ * nin_ * 2 + (long)width_ * width_;
    if (hipMalloc(&Wh_, sizeof(__bf16) * nw) != hipSuccess) return VES3D_MEMORY_ERROR;
  }
  act_ = mode;
  wdirty_ = true;
  // reallocate the batch buffers at the current capacity with the new
  // mode's operand buffers
  const long b = bcap_;
  bcap_ = 0;
  return b > 0 ? ensure_batch(b) : VES3D_SUCCESS;
}

int LearnedPC::refresh_bf16(hipStream_t st) {
  if (!wdirty_) return VES3D_SUCCESS;
  const long w = width_, n = nin_;
  k_to_bf16<<<gb(w * n), 256, 0, st>>>(P_ + oW1_, w * n, Wh_);
  V3D_CHECK(launch_status());
  k_to_bf16<<<gb(w * w), 256, 0, st>>>(P_ + oW2_, w * w, Wh_ + w * n);
  V3D_CHECK(launch_status());
  k_to_bf16<<<gb(n * w), 256, 0, st>>>(P_ + oW3_, n * w, Wh_ + w * n + w * w);
  V3D_CHECK(launch_status());
  wdirty_ = false;
  return VES3D_SUCCESS;
}

int LearnedPC::init(int nin, int width, long capacity, unsigned long long seed) {
  nin_ = nin;
  width_ = width;
  cap_ = capacity;
  seed_ = seed;
  oW1_ = 0;
  ob1_ = oW1_ + (long)width * nin;
  oW2_ = ob1_ + width;
  ob2_ = oW2_ + (long)width * width;
  oW3_ = ob2_ + width;
  ob3_ = oW3_ + (long)nin * width;
  np_ = ob3_ + nin;
  for (float **p : {&P_, &G_, &M_, &V_})
    if (hipMalloc(p, sizeof(float) * np_) != hipSuccess) return VES3D_MEMORY_ERROR;
  V3D_HIP(hipMemset(P_, 0, sizeof(float) * np_));
  V3D_HIP(hipMemset(M_, 0, sizeof(float) * np_));
  V3D_HIP(hipMemset(V_, 0, sizeof(float) * np_));
  // Glorot-uniform hidden layers; the output layer starts small so that
  // the untrained preconditioner is close to D.
  const float a1 = std::sqrt(6.0f / (nin + width)), a2 = std::sqrt(6.0f / (2.0f * width));
  k_init_uniform<<<gb((long)width * nin), 256>>>(P_ + oW1_, (long)width * nin, a1, seed * 3 + 1);
  k_init_uniform<<<gb((long)width * width), 256>>>(P_ + oW2_, (long)width * width, a2, seed * 3 + 2);
  k_init_uniform<<<gb((long)nin * width), 256>>>(P_ + oW3_, (long)nin * width, 0.01f * a1, seed * 3 + 3);
  V3D_CHECK(launch_status());
  if (cap_ > 0) {
    if (hipMalloc(&Zs_, sizeof(float) * cap_ * nin) != hipSuccess) return VES3D_MEMORY_ERROR;
    if (hipMalloc(&Ts_, sizeof(float) * cap_ * nin) != hipSuccess) return VES3D_MEMORY_ERROR;
  }
  if (hipMalloc(&lossbuf_, sizeof(double) * 4200) != hipSuccess) return VES3D_MEMORY_ERROR;
  if (hipMalloc(&evalbuf_, sizeof(double) * (2 + 2 * kEvalParts)) != hipSuccess)
    return VES3D_MEMORY_ERROR;
  return hip_status(hipDeviceSynchronize());
}

int LearnedPC::ensure_batch(long B) {
  if (B <= bcap_) return VES3D_SUCCESS;
  for (float *p : {A1_, H1_, A2_, H2_, Y_, dY_, dA2_, dA1_, Zb_, Tb_, Zneg_, Yneg_})
    if (p) (void)hipFree(p);
  if (scale_) (void)hipFree(scale_);
  const long w = width_, n = nin_;
  float **wb[] = {&A1_, &H1_, &A2_, &H2_, &dA2_, &dA1_};
  for (float **p : wb)
    if (hipMalloc(p, sizeof(float) * B * w) != hipSuccess) return VES3D_MEMORY_ERROR;
  float **nb[] = {&Y_, &dY_, &Zb_, &Tb_, &Zneg_, &Yneg_};
  for (float **p : nb)
    if (hipMalloc(p, sizeof(float) * B * n) != hipSuccess) return VES3D_MEMORY_ERROR;
  if (hipMalloc(&scale_, sizeof(double) * B) != hipSuccess) return VES3D_MEMORY_ERROR;
  for (__bf16 *p : {Zh_, H1h_, H2h_, dYh_, dA2h_, dA1h_})
    if (p) (void)hipFree(p);
  Zh_ = H1h_ = H2h_ = dYh_ = dA2h_ = dA1h_ = nullptr;
  if (act_ == 1) {
    for (__bf16 **p : {&H1h_, &H2h_, &dA2h_, &dA1h_})
      if (hipMalloc(p, sizeof(__bf16) * B * w) != hipSuccess) return VES3D_MEMORY_ERROR;
    for (__bf16 **p : {&Zh_, &dYh_})
      if (hipMalloc(p, sizeof(__bf16) * B * n) != hipSuccess) return VES3D_MEMORY_ERROR;
  }
  bcap_ = B;
  return VES3D_SUCCESS;
}

// C[M][N] = A B with A(r,k) = A[r ars + k acs], B(k,c) = B[k brs + c bcs]
int LearnedPC::gemm(int M, int N, int K, const float *A, long ars, long acs, const float *Bm,
                    long brs, long bcs, float *C, hipStream_t st) {
  auto key = std::make_tuple(M, N, K, ars, acs, brs, bcs);
  GemmDesc h{};
  h.M = M;
  h.N = N;
  h.K = K;
  h.a_rs = ars;
  h.a_cs = acs;
  h.b_rs = brs;
  h.b_cs = bcs;
  h.c_rs = N;
  h.c_cs = 1;
  GemmDesc *d = nullptr;
  auto it = descs_.find(key);
  if (it != descs_.end()) {
    d = it->second;
  } else {
    d = upload_descs(&h, 1);
    if (!d) return VES3D_MEMORY_ERROR;
    descs_[key] = d;
  }
  return gemm_grouped<float>(d, 1, 1, M, N, A, Bm, C, 1.0f, 0.0f, st, gemm_pick_tile(&h, 1));
}

int LearnedPC::forward(const float *Z, long B, float *Y, hipStream_t st) {
  V3D_CHECK(ensure_batch(B));
  const int w = width_, n = nin_;
  if (act_ == 1) {
    // BF16 operands (weights rounded from the FP32 master copy), FP32
    // accumulation, FP32 pre-activations and output
    V3D_CHECK(refresh_bf16(st));
    const __bf16 *W1h = Wh_, *W2h = Wh_ + (long)w * n, *W3h = W2h + (long)w * w;
    k_to_bf16<<<gb(B * n), 256, 0, st>>>(Z, B * n, Zh_);
    V3D_CHECK(launch_status());
    V3D_CHECK((gemm_bf16<true, true>((int)B, w, n, Zh_, n, W1h, n, A1_, w, st)));
    k_bias_gelu_bf<<<gb(B * w), 256, 0, st>>>(A1_, P_ + ob1_, B, w, H1h_);
    V3D_CHECK(launch_status());
    V3D_CHECK((gemm_bf16<true, true>((int)B, w, w, H1h_, w, W2h, w, A2_, w, st)));
    k_bias_gelu_bf<<<gb(B * w), 256, 0, st>>>(A2_, P_ + ob2_, B, w, H2h_);
    V3D_CHECK(launch_status());
    V3D_CHECK((gemm_bf16<true, true>((int)B, n, w, H2h_, w, W3h, w, Y, n, st)));
    k_add_bias<<<gb(B * n), 256, 0, st>>>(Y, P_ + ob3_, B, n);
    return launch_status();
  }
  V3D_CHECK(gemm((int)B, w, n, Z, n, 1, P_ + oW1_, 1, n, A1_, st));   // Z W1^T
  k_bias_gelu<<<gb(B * w), 256, 0, st>>>(A1_, P_ + ob1_, B, w, H1_);
  V3D_CHECK(launch_status());
  V3D_CHECK(gemm((int)B, w, w, H1_, w, 1, P_ + oW2_, 1, w, A2_, st));  // H1 W2^T
  k_bias_gelu<<<gb(B * w), 256, 0, st>>>(A2_, P_ + ob2_, B, w, H2_);
  V3D_CHECK(launch_status());
  V3D_CHECK(gemm((int)B, n, w, H2_, w, 1, P_ + oW3_, 1, w, Y, st));    // H2 W3^T
  k_add_bias<<<gb(B * n), 256, 0, st>>>(Y, P_ + ob3_, B, n);
  return launch_status();
}

int LearnedPC::loss_grad(const float *Z, const float *T, long B, double *loss_host,
                         hipStream_t st) {
  V3D_CHECK(forward(Z, B, Y_, st));
  const int w = width_, n = nin_;
  const long tot = B * n;
  const int nb = (int)std::min<long>(4096, std::max<long>(1, (tot + 255) / 256));
  k_diff_loss<<<nb, 256, 0, st>>>(Y_, T, tot, 1.0f / (float)tot, dY_, lossbuf_);
  V3D_CHECK(launch_status());
  k_sum_parts<<<1, 64, 0, st>>>(lossbuf_, nb, 0.5 / (double)tot, lossbuf_ + 4100);
  V3D_CHECK(launch_status());
  if (act_ == 1) {
    // BF16 operands of every product, FP32 gradients and bias sums
    const __bf16 *W2h = Wh_ + (long)w * n, *W3h = W2h + (long)w * w;
    const int Bi = (int)B;
    k_to_bf16<<<gb(tot), 256, 0, st>>>(dY_, tot, dYh_);
    V3D_CHECK(launch_status());
    V3D_CHECK((gemm_bf16<false, false>(n, w, Bi, dYh_, n, H2h_, w, G_ + oW3_, w, st)));
    k_colsum<<<gb(n), 256, 0, st>>>(dY_, B, n, G_ + ob3_);
    V3D_CHECK(launch_status());
    V3D_CHECK((gemm_bf16<true, false>(Bi, w, n, dYh_, n, W3h, w, dA2_, w, st)));
    k_gelu_back_bf<<<gb(B * w), 256, 0, st>>>(dA2_, A2_, B * w, dA2h_);
    V3D_CHECK(launch_status());
    V3D_CHECK((gemm_bf16<false, false>(w, w, Bi, dA2h_, w, H1h_, w, G_ + oW2_, w, st)));
    k_colsum<<<gb(w), 256, 0, st>>>(dA2_, B, w, G_ + ob2_);
    V3D_CHECK(launch_status());
    V3D_CHECK((gemm_bf16<true, false>(Bi, w, w, dA2h_, w, W2h, w, dA1_, w, st)));
    k_gelu_back_bf<<<gb(B * w), 256, 0, st>>>(dA1_, A1_, B * w, dA1h_);
    V3D_CHECK(launch_status());
    V3D_CHECK((gemm_bf16<false, false>(w, n, Bi, dA1h_, w, Zh_, n, G_ + oW1_, n, st)));
    k_colsum<<<gb(w), 256, 0, st>>>(dA1_, B, w, G_ + ob1_);
    V3D_CHECK(launch_status());
    if (loss_host) {
      V3D_HIP(hipMemcpyAsync(loss_host, lossbuf_ + 4100, sizeof(double), hipMemcpyDeviceToHost,
                             st));
      V3D_HIP(hipStreamSynchronize(st));
    }
    return VES3D_SUCCESS;
  }
  // layer 3
  V3D_CHECK(gemm(n, w, (int)B, dY_, 1, n, H2_, w, 1, G_ + oW3_, st));  // dY^T H2
  k_colsum<<<gb(n), 256, 0, st>>>(dY_, B, n, G_ + ob3_);
  V3D_CHECK(launch_status());
  V3D_CHECK(gemm((int)B, w, n, dY_, n, 1, P_ + oW3_, w, 1, dA2_, st));  // dY W3
  k_gelu_back<<<gb(B * w), 256, 0, st>>>(dA2_, A2_, B * w);
  V3D_CHECK(launch_status());
  // layer 2
  V3D_CHECK(gemm(w, w, (int)B, dA2_, 1, w, H1_, w, 1, G_ + oW2_, st));
  k_colsum<<<gb(w), 256, 0, st>>>(dA2_, B, w, G_ + ob2_);
  V3D_CHECK(launch_status());
  V3D_CHECK(gemm((int)B, w, w, dA2_, w, 1, P_ + oW2_, w, 1, dA1_, st));
  k_gelu_back<<<gb(B * w), 256, 0, st>>>(dA1_, A1_, B * w);
  V3D_CHECK(launch_status());
  // layer 1
  V3D_CHECK(gemm(w, n, (int)B, dA1_, 1, w, Z, n, 1, G_ + oW1_, st));
  k_colsum<<<gb(w), 256, 0, st>>>(dA1_, B, w, G_ + ob1_);
  V3D_CHECK(launch_status());
  if (loss_host) {
    V3D_HIP(hipMemcpyAsync(loss_host, lossbuf_ + 4100, sizeof(double), hipMemcpyDeviceToHost, st));
    V3D_HIP(hipStreamSynchronize(st));
  }
  return VES3D_SUCCESS;
}

int LearnedPC::reset_optimizer() {
  V3D_HIP(hipMemset(M_, 0, sizeof(float) * np_));
  V3D_HIP(hipMemset(V_, 0, sizeof(float) * np_));
  t_ = 0;
  return VES3D_SUCCESS;
}

int LearnedPC::adam(float lr, hipStream_t st) {
  ++t_;
  const float c1 = 1.0f - std::pow(0.9f, (float)t_), c2 = 1.0f - std::pow(0.999f, (float)t_);
  k_adam<<<gb(np_), 256, 0, st>>>(P_, G_, M_, V_, np_, lr, c1, c2, wd_);
  wdirty_ = true;
  return launch_status();
}

int LearnedPC::subset_loss(const int *idx, long m, double *loss_host, double *target_ms_host,
                            hipStream_t st) {
  const long B = std::min<long>(std::max<long>(bcap_, 1), m);
  V3D_CHECK(ensure_batch(B));
  double *acc = evalbuf_, *part = evalbuf_ + 2;
  V3D_HIP(hipMemsetAsync(acc, 0, 2 * sizeof(double), st));
  for (long r0 = 0; r0 < m; r0 += B) {
    const long cnt = std::min(B, m - r0);
    k_gather_batch<<<gb(cnt * nin_), 256, 0, st>>>(Zs_, Ts_, idx + r0, cnt, nin_, Zb_, Tb_);
    V3D_CHECK(launch_status());
    V3D_CHECK(odd_forward(Zb_, cnt, Y_, st));
    const long tot = cnt * nin_;
    const int nb = (int)std::min<long>(kEvalParts, std::max<long>(1, (tot + 255) / 256));
    k_sq_sums<<<nb, 256, 0, st>>>(Y_, Tb_, tot, part);
    V3D_CHECK(launch_status());
    k_acc_pairs<<<1, 64, 0, st>>>(part, nb, acc);
    V3D_CHECK(launch_status());
  }
  double h[2];
  V3D_HIP(hipMemcpyAsync(h, acc, sizeof(h), hipMemcpyDeviceToHost, st));
  V3D_HIP(hipStreamSynchronize(st));
  const double n = (double)m * nin_;
  if (loss_host) *loss_host = 0.5 * h[0] / n;
  if (target_ms_host) *target_ms_host = h[1] / n;
  return VES3D_SUCCESS;
}

int LearnedPC::dataset_loss(long have, double *loss_host, double *target_ms_host,
                            hipStream_t st) {
  if (!perm_ && hipMalloc(&perm_, sizeof(int) * cap_) != hipSuccess) return VES3D_MEMORY_ERROR;
  std::vector<int> all(have);
  for (long i = 0; i < have; ++i) all[i] = (int)i;
  V3D_HIP(hipStreamSynchronize(st));
  V3D_HIP(hipMemcpy(perm_, all.data(), sizeof(int) * have, hipMemcpyHostToDevice));
  return subset_loss(perm_, have, loss_host, target_ms_host, st);
}

int LearnedPC::train(int epochs, int batch, float lr, double *last_loss, hipStream_t st) {
  const long have = std::min(count_, cap_);
  if (have < 2 || batch < 1 || epochs < 1) return VES3D_INVALID_PARAMETER;
  // a fixed tenth of the samples is held out; the kept parameters are those
  // with the lowest held-out loss (evaluated after every epoch)
  const long nval = std::max<long>(1, have / 10), ntr = have - nval;
  const long B = std::min<long>(batch, ntr);
  V3D_CHECK(ensure_batch(B));
  if (!perm_ && hipMalloc(&perm_, sizeof(int) * cap_) != hipSuccess) return VES3D_MEMORY_ERROR;
  if (!vidx_ && hipMalloc(&vidx_, sizeof(int) * cap_) != hipSuccess) return VES3D_MEMORY_ERROR;
  if (!Pbest_ && hipMalloc(&Pbest_, sizeof(float) * np_) != hipSuccess) return VES3D_MEMORY_ERROR;
  std::mt19937_64 split(seed_ + 777);
  std::vector<int> all(have);
  for (long i = 0; i < have; ++i) all[i] = (int)i;
  std::shuffle(all.begin(), all.end(), split);
  std::vector<int> tr(all.begin() + nval, all.end());
  V3D_HIP(hipStreamSynchronize(st));
  V3D_HIP(hipMemcpy(vidx_, all.data(), sizeof(int) * nval, hipMemcpyHostToDevice));
  std::mt19937_64 rng(seed_ + 12345 + (unsigned long long)t_);
  const long per_epoch = std::max<long>(1, ntr / B);
  const long total = epochs * per_epoch;
  // Adam over shuffled minibatches with a cosine learning-rate decay
  double best = 0;
  V3D_CHECK(subset_loss(vidx_, nval, &best, nullptr, st));
  V3D_HIP(hipMemcpyAsync(Pbest_, P_, sizeof(float) * np_, hipMemcpyDeviceToDevice, st));
  long k = 0;
  for (int ep = 0; ep < epochs; ++ep) {
    std::shuffle(tr.begin(), tr.end(), rng);
    V3D_HIP(hipStreamSynchronize(st));
    V3D_HIP(hipMemcpy(perm_, tr.data(), sizeof(int) * ntr, hipMemcpyHostToDevice));
    for (long b = 0; b < per_epoch; ++b, ++k) {
      k_gather_batch<<<gb(B * nin_), 256, 0, st>>>(Zs_, Ts_, perm_ + b * B, B, nin_, Zb_, Tb_);
      V3D_CHECK(launch_status());
      V3D_CHECK(loss_grad(Zb_, Tb_, B, nullptr, st));
      const double c = 0.5 * (1.0 + std::cos(M_PI * (double)k / (double)total));
      V3D_CHECK(adam((float)(lr * c), st));
    }
    double l = 0;
    V3D_CHECK(subset_loss(vidx_, nval, &l, nullptr, st));
    if (l < best) {
      best = l;
      best_epoch_ = ep + 1;
      V3D_HIP(hipMemcpyAsync(Pbest_, P_, sizeof(float) * np_, hipMemcpyDeviceToDevice, st));
    }
  }
  V3D_HIP(hipMemcpyAsync(P_, Pbest_, sizeof(float) * np_, hipMemcpyDeviceToDevice, st));
  wdirty_ = true;
  if (last_loss) *last_loss = best;
  trained_ = true;
  return hip_status(hipStreamSynchronize(st));
}

int LearnedPC::apply_inplace(double *y, long nv, int p, int tcut, const int2 *cinfo,
                             const double *frames, hipStream_t st) {
  const int nc = p * (p + 2);
  if (4 * nc != nin_) return VES3D_INVALID_PARAMETER;
  V3D_CHECK(ensure_batch(nv));
  k_row_norms<<<(unsigned)nv, 256, 0, st>>>(y, nv, nc, cinfo, scale_);
  V3D_CHECK(launch_status());
  k_gather_rows<<<gb(nv * nin_), 256, 0, st>>>(y, nv, nc, cinfo, scale_, frames, Zb_);
  V3D_CHECK(launch_status());
  V3D_CHECK(odd_forward(Zb_, nv, Y_, st));
  k_scatter_add<<<gb(nv * nin_), 256, 0, st>>>(Y_, nv, nc, p, tcut, cinfo, scale_, frames,
                                                y);
  return launch_status();
}

int LearnedPC::odd_forward(const float *Z, long B, float *Y, hipStream_t st) {
  V3D_CHECK(ensure_batch(B));
  const long n = B * nin_;
  V3D_CHECK(forward(Z, B, Y, st));
  k_negate<<<gb(n), 256, 0, st>>>(Z, n, Zneg_);
  V3D_CHECK(launch_status());
  V3D_CHECK(forward(Zneg_, B, Yneg_, st));
  k_odd_part<<<gb(n), 256, 0, st>>>(Y, Yneg_, n);
  return launch_status();
}

int LearnedPC::record(const double *Dr, const double *e, long nv, int p, const int2 *cinfo,
                      const double *frames, hipStream_t st) {
  const int nc = p * (p + 2);
  if (4 * nc != nin_ || cap_ < nv) return VES3D_INVALID_PARAMETER;
  V3D_CHECK(ensure_batch(nv));
  k_row_norms<<<(unsigned)nv, 256, 0, st>>>(Dr, nv, nc, cinfo, scale_);
  V3D_CHECK(launch_status());
  k_record<<<gb(nv * nin_), 256, 0, st>>>(Dr, e, nv, nc, cinfo, scale_, frames, head_, cap_,
                                           Zs_, Ts_);
  V3D_CHECK(launch_status());
  head_ = (head_ + nv) % cap_;
  count_ += nv;
  return VES3D_SUCCESS;
}

}  // namespace v3d

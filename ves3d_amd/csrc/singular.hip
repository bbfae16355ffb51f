// singular.hip -- rotations of SH coefficients to every grid point
// (RotateAll / RotateTranspose, src/SphericalHarmonics.cc:362-541), the
// singular self-interaction operator build (StokesSingularInteg_,
// :952-1180) and its per-vesicle application (the self term of
// StokesVelocity::operator(), src/StokesVelocity.cc:385-447).
//
// Rotation: the per-degree block products Bo[l] = Bi[l] * Mr[t][l] run as one
// grouped GEMM launch (descriptor per (t, l), batch over vesicles) on a
// degree-major copy of the coefficients; the longitude phase and the
// degree-major gather/scatter are separate elementwise kernels.
//
// Singular build, per chunk of vesicles:
//   Grid2SHC -> RotateAll -> SHC2Grid(p0 -> p1, +derivatives) & SHC2Pole
//   -> pole kernel (FP64) -> SHC2GridTranspose(p1 -> p0) -> RotateTranspose
//   -> rearrange -> Grid2SHC = [Nves][3 Ncoef][3 Ncoef]
// The FP64 build of p1 = 2 p0 <= 32 fuses the pole kernel with
// SHC2GridTranspose and RotateTranspose's gather (k_pole_transpose), and the
// rearrangement with the last Grid2SHC (k_rearrange_g2s).  The FP32 form runs
// every GEMM stage in FP32 and keeps the pole kernel (singular weights, 1/r
// near the rotated pole) in FP64.  Both write the operator's (a, b) and
// (b, a) blocks equal (the rows of one symmetric tensor component), which
// the step's self apply relies on (k_self_apply_sym).
#include <algorithm>
#include <atomic>
#include <cmath>
#include <map>
#include <mutex>
#include <tuple>
#include <type_traits>
#include <vector>

#include "runtime.hpp"
#include "sht.hpp"

namespace v3d {

namespace {

std::mutex g_mu;
// The build's scratch (kWsSing0) is one per device: each build waits on its
// stream for the device's previous build, wherever that ran
struct BuildOrder {
  hipEvent_t done = nullptr;
  bool recorded = false;
};
std::map<int, BuildOrder> g_build_order;
// 1: the singular build's pole kernel runs fused with the transposed
// synthesis (k_pole_transpose) where its shapes allow; 0: as separate
// kernel and grouped GEMMs (ves3d_hip_set_pole_kernel)
std::atomic<int> g_pole_fused{1};
struct RotPlan {
  GemmDesc *d;
  int n, tile;
};
std::map<std::tuple<int, int, int, int, long>, RotPlan> g_rplans;

// Grouped GEMM plan of the per-degree rotation products.
// transpose=false: RotateAll, rows (j, d) of vesicle v, output slab t.
// transpose=true : RotateTranspose, rows (j, d) of (v, t), Mr read transposed.
const RotPlan *rot_plan(int p, int dof, bool transpose,
                                           const RotMatsHost &h) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return nullptr;
  std::lock_guard<std::mutex> lk(g_mu);
  auto key = std::make_tuple(dev, p, dof, (int)transpose, 0L);
  auto it = g_rplans.find(key);
  if (it != g_rplans.end()) return &it->second;
  const long Ncoef = sh_ncoef(p);
  const long slab = (long)p * dof * Ncoef;  // one (v, t) slab
  std::vector<GemmDesc> v;
  for (int t = 0; t <= p; ++t)
    for (int l = 0; l <= p; ++l) {
      GemmDesc d{};
      const int len = h.len[l];
      d.M = p * dof;
      d.N = len;
      d.K = len;
      d.a_rs = Ncoef;
      d.a_cs = 1;
      d.b_off = h.off[t * (p + 1) + l];
      d.b_rs = transpose ? 1 : len;
      d.b_cs = transpose ? len : 1;
      d.b_bs = 0;
      d.c_rs = Ncoef;
      d.c_cs = 1;
      if (!transpose) {
        d.a_off = h.doff[l];
        d.a_bs = slab;
        d.c_off = t * slab + h.doff[l];
        d.c_bs = (p + 1) * slab;
      } else {
        d.a_off = t * slab + h.doff[l];
        d.a_bs = (p + 1) * slab;
        d.c_off = t * slab + h.doff[l];
        d.c_bs = (p + 1) * slab;
      }
      v.push_back(d);
    }
  GemmDesc *dd = upload_descs(v.data(), (int)v.size());
  if (!dd) return nullptr;
  return &(g_rplans[key] = RotPlan{dd, (int)v.size(), gemm_pick_tile(v.data(), (int)v.size())});
}

template <typename T>
const T *rot_ptr(const RotMatsDev *r);
template <>
const double *rot_ptr<double>(const RotMatsDev *r) { return r->d64; }
template <>
const float *rot_ptr<float>(const RotMatsDev *r) { return r->d32; }

// The four coefficient permutation kernels below run one wavefront per row
// of Ncoef coefficients (four rows per 256-thread block): the row's indices
// are decomposed once per wavefront, lanes stride over the row, and the
// permuted reads stay inside one row (L1-resident).

// B0p[v][j][d][pos] = phase_j(S0[v][d])[perm[pos]] with exp(-i k j pi / p)
// (src/SphericalHarmonics.cc:398-418); phase[j][k] = (cos, sin)(k j pi / p).
// Rows: (v, j, d).
template <typename T>
__global__ __launch_bounds__(256) void k_rot_phase_gather(
    const T *__restrict__ S0, long rows, int p, int dof, int ncoef,
    const int *__restrict__ perm, const int *__restrict__ partner,
    const int *__restrict__ order, const int *__restrict__ is_sin,
    const double *__restrict__ phase, T *__restrict__ out) {
  const long row = (long)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (row >= rows) return;
  const int d = (int)(row % dof);
  const long r = row / dof;
  const int j = (int)(r % p);
  const long v = r / p;
  const T *s = S0 + (v * dof + d) * ncoef;
  T *o = out + row * ncoef;
  for (int pos = threadIdx.x & 63; pos < ncoef; pos += 64) {
    const int c = perm[pos];
    const int k = order[c];
    T val;
    if (k == 0 || k == p) {
      val = s[c];
    } else {
      const bool sn = is_sin[c];
      const T x0 = s[sn ? partner[c] : c], x1 = s[sn ? c : partner[c]];
      const double cv = phase[2 * (j * (p + 1) + k)], sv = phase[2 * (j * (p + 1) + k) + 1];
      const T r0 = (T)cv, r1 = (T)(-sv);
      val = sn ? x0 * r1 + x1 * r0 : x0 * r0 - x1 * r1;
    }
    o[pos] = val;
  }
}

// out[row][c] = in[row][iperm[c]]  (degree-major -> coefficient order)
template <typename T>
__global__ __launch_bounds__(256) void k_scatter_coef(const T *__restrict__ in, long rows,
                                                      int ncoef, const int *__restrict__ iperm,
                                                      T *__restrict__ out) {
  const long row = (long)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (row >= rows) return;
  const T *s = in + row * ncoef;
  T *o = out + row * ncoef;
  for (int c = threadIdx.x & 63; c < ncoef; c += 64) o[c] = s[iperm[c]];
}

// out[row][pos] = in[row][perm[pos]]  (coefficient order -> degree-major)
template <typename T>
__global__ __launch_bounds__(256) void k_gather_coef(const T *__restrict__ in, long rows,
                                                     int ncoef, const int *__restrict__ perm,
                                                     T *__restrict__ out) {
  const long row = (long)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (row >= rows) return;
  const T *s = in + row * ncoef;
  T *o = out + row * ncoef;
  for (int pos = threadIdx.x & 63; pos < ncoef; pos += 64) o[pos] = s[perm[pos]];
}

// out[v][t][j][d][c] = phase_j^+(in)[c] reading the degree-major input
// (src/SphericalHarmonics.cc:511-537).  Rows: (v, t, j, d).  tab[c] holds
// c's two degree-major sources, its order and its kind (copy, cos, sin), so
// each output takes one 16-byte table read before its data reads.
template <typename T, bool STAGED>
__global__ __launch_bounds__(256) void k_rot_scatter_phase(const T *__restrict__ in, long rows,
                                                           int p, int dof, int ncoef,
                                                           const int4 *__restrict__ tab,
                                                           const double *__restrict__ phase,
                                                           T *__restrict__ out) {
  // STAGED: the block's four input rows go through LDS with contiguous
  // reads, and the permuted reads hit LDS (rows up to kScatterStageBytes)
  extern __shared__ unsigned char rsp_smem[];
  const long row = (long)blockIdx.x * 4 + (threadIdx.x >> 6);
  const bool live = row < rows;
  const T *s = in + row * ncoef;
  if constexpr (STAGED) {
    T *srow = (T *)rsp_smem + (threadIdx.x >> 6) * ncoef;
    if (live)
      for (int c = threadIdx.x & 63; c < ncoef; c += 64) srow[c] = s[c];
    __syncthreads();
    s = srow;
  }
  if (!live) return;
  const int j = (int)((row / dof) % p);
  const double *ph = phase + 2L * j * (p + 1);
  T *o = out + row * ncoef;
  for (int c = threadIdx.x & 63; c < ncoef; c += 64) {
    const int4 e = tab[c];
    T val;
    if (e.w == 0) {
      val = s[e.x];
    } else {
      const T x0 = s[e.x], x1 = s[e.y];
      const T r0 = (T)ph[2 * e.z], r1 = (T)ph[2 * e.z + 1];
      val = e.w == 2 ? x0 * r1 + x1 * r0 : x0 * r0 - x1 * r1;
    }
    o[c] = val;
  }
}

// Largest LDS staging of four rows (the rotation scatter at p <= 31 in FP64).
constexpr size_t kScatterStageBytes = 32768;

// Blocks of four rows.
inline unsigned row_blocks(long rows) { return (unsigned)((rows + 3) / 4); }

inline unsigned grid_for(long n) {
  long b = (n + 255) / 256;
  if (b > 65536L * 16) b = 65536L * 16;
  return (unsigned)(b < 1 ? 1 : b);
}

}  // namespace

// RotateAll: S [Nv*dof][Ncoef] -> out [Nv][p+1][p][dof][Ncoef]; work: Nv*p*dof*Ncoef
// + Nv*(p+1)*p*dof*Ncoef elements.
template <typename T>
int rotate_all(const T *S, long Nv, int p, int dof, T *out, T *work, hipStream_t st) {
  const RotMatsDev *r = rot_mats(p);
  if (!r) return VES3D_MEMORY_ERROR;
  const auto *pl = rot_plan(p, dof, false, *r->h);
  if (!pl) return VES3D_MEMORY_ERROR;
  const int nc = sh_ncoef(p);
  T *B0p = work;
  T *S1p = work + Nv * p * dof * (long)nc;
  const long rows0 = Nv * p * (long)dof;
  const long rows = Nv * (p + 1) * (long)p * dof;
  if (rows / 4 >= 0x7fffffffL) return VES3D_INVALID_PARAMETER;
  k_rot_phase_gather<T><<<row_blocks(rows0), 256, 0, st>>>(
      S, rows0, p, dof, nc, r->perm, r->partner, r->order, r->is_sin, r->phase, B0p);
  V3D_CHECK(launch_status());
  V3D_CHECK(gemm_grouped<T>(pl->d, pl->n, (int)Nv, p * dof, 2 * p, B0p,
                            rot_ptr<T>(r), S1p, T(1), T(0), st, pl->tile));
  k_scatter_coef<T><<<row_blocks(rows), 256, 0, st>>>(S1p, rows, nc, r->iperm, out);
  return launch_status();
}

// RotateTranspose: in/out [Nv][p+1][p][dof][Ncoef]; work: 2 * that size
// (that size when gathered: the input rows are already in degree-major order).
template <typename T>
int rotate_transpose(const T *in, long Nv, int p, int dof, T *out, T *work, hipStream_t st,
                     bool gathered) {
  const RotMatsDev *r = rot_mats(p);
  if (!r) return VES3D_MEMORY_ERROR;
  const auto *pl = rot_plan(p, dof, true, *r->h);
  if (!pl) return VES3D_MEMORY_ERROR;
  const int nc = sh_ncoef(p);
  const long rows = Nv * (p + 1) * (long)p * dof;
  if (rows / 4 >= 0x7fffffffL) return VES3D_INVALID_PARAMETER;
  const T *Ap = in;
  T *Bp = work;
  if (!gathered) {
    k_gather_coef<T><<<row_blocks(rows), 256, 0, st>>>(in, rows, nc, r->perm, work);
    V3D_CHECK(launch_status());
    Ap = work;
    Bp = work + rows * nc;
  }
  V3D_CHECK(gemm_grouped<T>(pl->d, pl->n, (int)Nv, p * dof, 2 * p, Ap,
                            rot_ptr<T>(r), Bp, T(1), T(0), st, pl->tile));
  const size_t stage = 4 * nc * sizeof(T);
  if (stage <= kScatterStageBytes)
    k_rot_scatter_phase<T, true><<<row_blocks(rows), 256, stage, st>>>(
        Bp, rows, p, dof, nc, (const int4 *)r->scatter, r->phase, out);
  else
    k_rot_scatter_phase<T, false><<<row_blocks(rows), 256, 0, st>>>(
        Bp, rows, p, dof, nc, (const int4 *)r->scatter, r->phase, out);
  return launch_status();
}

// ------------------------------------------------------------- pole kernel
// For each rotation r, pole t (0: x=-1, 1: x=+1) and upsampled source s:
// the 6 symmetric tensor components of the SL (and/or DL) kernel times the
// singular quadrature weight (src/SphericalHarmonics.cc:970-1073).  One
// thread per (r, s) serves both poles, so the rotated geometry (X, X_theta,
// X_phi: 72 bytes in FP64) is read once for the two 48-byte outputs.
template <typename T>
__global__ void k_pole_kernel(const T *__restrict__ X, const T *__restrict__ Xt,
                              const T *__restrict__ Xp, const T *__restrict__ trg,
                              const double *__restrict__ qw, long nrot, int p1,
                              T *__restrict__ SL0, T *__restrict__ DL0) {
  const int M1 = 2 * p1 * (p1 + 1);
  const long total = nrot * (long)M1;
  const double eps = 2.220446049250313e-16;
  const double csl = 1.0 / (8.0 * kPi), cdl = 3.0 / (4.0 * kPi);
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < total;
       i += (long)gridDim.x * blockDim.x) {
    const int s = (int)(i % M1);
    const long r = i / M1;
    const int j0 = s / (2 * p1);
    const long b = r * 3 * (long)M1 + s;
    const double x0 = X[b], y0 = X[b + M1], z0 = X[b + 2 * M1];
    const double xt = Xt[b], yt = Xt[b + M1], zt = Xt[b + 2 * M1];
    const double xp = Xp[b], yp = Xp[b + M1], zp = Xp[b + 2 * M1];
    const double nx = yt * zp - zt * yp, ny = zt * xp - xt * zp, nz = xt * yp - yt * xp;
    const double area = SL0 ? sqrt(nx * nx + ny * ny + nz * nz) : 0.0;
#pragma unroll
    for (int t = 0; t < 2; ++t) {
      const double w = qw[t ? j0 : p1 - j0];
      const double dx = trg[(r * 3 + 0) * 2 + t] - x0, dy = trg[(r * 3 + 1) * 2 + t] - y0,
                   dz = trg[(r * 3 + 2) * 2 + t] - z0;
      const double r2 = dx * dx + dy * dy + dz * dz;
      double rinv = 1.0 / sqrt(r2);
      if (r2 <= eps) rinv = 0;
      const double rinv2 = rinv * rinv;
      const double dd[6] = {dx * dx, dx * dy, dx * dz, dy * dy, dy * dz, dz * dz};
      const long o = ((2 * r + t) * 6) * (long)M1 + s;
      if (DL0) {
        const double rn5 = cdl * w * (nx * dx + ny * dy + nz * dz) * rinv2 * rinv2 * rinv;
#pragma unroll
        for (int c = 0; c < 6; ++c) DL0[o + c * (long)M1] = (T)(dd[c] * rn5);
      }
      if (SL0) {
        const double ar = csl * w * area * rinv;
        const double ar2 = ar * rinv2;
#pragma unroll
        for (int c = 0; c < 6; ++c)
          SL0[o + c * (long)M1] = (T)(dd[c] * ar2 + ((c == 0 || c == 3 || c == 5) ? ar : 0.0));
      }
    }
  }
}

// The pole kernel fused with SHC2GridTranspose(p1 -> p0) of its values:
// K1[r][t*6+c][dpos[k]] = sum over the grid (lat, lon) of the kernel value of
// rotation r, pole t, tensor component c at (lat, lon), times MatFourier(p0,
// p1)[jj][lon] (the Fourier stage) and MatLegendre(p0, p1)[m][n][lat] (the
// Legendre stage), k = blockoff(p0, jj) + n, m = (jj + 1) / 2 -- the products
// the two grouped GEMMs of sht_shc2grid_transpose form, without the kernel
// values (12 functions on the upsampled grid per rotation) or the Fourier
// stage's output passing through HBM.  dpos (RotateTranspose's degree-major
// position of each coefficient) lays the rows out as RotateTranspose's
// per-degree products read them, so its gather pass is not needed.
// Persistent workgroups walk the rotations r = blockIdx.x + i gridDim.x and,
// within a rotation, the latitude rows in passes of R rows:
//   1. thread (row, lon, t) evaluates pole t's 6 kernel values at (lat, lon)
//      into Kv[row][lon][t * 6 + q] (kPtKs doubles per point, an odd stride);
//   2. wave (row, tile) forms the 16 x 16 tile T[row][c][16 tile + j] =
//      sum_lon Kv[row][lon][c] Mf[lon][16 tile + j] on v_mfma_f64_16x16x4f64
//      (lane l: A(l % 16, l / 16), B(l / 16, l % 16), results at rows
//      l / 16 + 4 i, column l % 16; rows c >= 12 read as zero), its B operand
//      (the wave's 16 columns of MatFourier) in registers for the whole kernel;
//   3. thread (c, jj) adds T[row][c][jj] times the rows' Legendre values into
//      its p0 + 1 - m accumulators (FP64 registers).
// Two barriers per pass; each kernel-value thread loads the geometry of its
// point of the next pass while the current pass runs.  At a rotation's end
// the accumulators go through LDS (over Kv and T) in the output layout and
// leave as whole rows, as many rows per round as that space holds.  The Legendre blocks and the weights sit in LDS for the workgroup's
// lifetime.
constexpr int kPtKs = 13;

// LDS doubles of k_pole_transpose: Kv, T, the weights, the Legendre blocks
// and LMAX - 1 zero rows after them (the branch-free degree loop's reads
// past a block's end).
inline long pole_lds_doubles(int R, int p0, int p1, int lmax) {
  const long L = 2 * p1, Nlat = p1 + 1, FP = 16 * ((2L * p0 + 15) / 16);
  return R * L * kPtKs + R * 12 * FP + Nlat + (long)(p0 + 1) * (p0 + 2) / 2 * Nlat +
         (lmax - 1) * Nlat;
}

template <int LMAX, int P1>
__global__ __launch_bounds__(384) void k_pole_transpose(
    const double *__restrict__ X, const double *__restrict__ Xt, const double *__restrict__ Xp,
    const double *__restrict__ trg, const double *__restrict__ qw, long nrot, int p0, int R,
    const double *__restrict__ mft, const double *__restrict__ mlt, int layer,
    const int *__restrict__ dpos, double *__restrict__ K1) {
  typedef double v4d __attribute__((ext_vector_type(4)));
  // FOLD (p1 = 32): longitudes lon and lon + L/2 differ by pi, where the
  // Fourier functions of order m change sign by (-1)^m, so a wave forms the
  // 16 columns of one parity of m from Kv[lon] +- Kv[lon + L/2], lon < L/2:
  // half the MFMA steps.  Wave tiles: the two parities (FOLD), else 16
  // consecutive columns.
  constexpr bool FOLD = P1 == 32;
  constexpr int L = 2 * P1, Nlat = P1 + 1, M1 = L * Nlat, KS = FOLD ? L / 8 : L / 4;
  extern __shared__ double lds[];
  const int F = 2 * p0, FP = 16 * ((F + 15) / 16), NT = FOLD ? 2 : FP / 16,
            Nc = p0 * (p0 + 2);
  const int nmlt = (p0 + 1) * (p0 + 2) / 2 * Nlat;
  double *Kv = lds;                 // [R][L][kPtKs]
  double *Ts = Kv + R * L * kPtKs;  // [R][12][FP]
  double *Qw = Ts + R * 12 * FP;    // [Nlat]
  double *Ml = Qw + Nlat;           // blocks m = 0..p0: [(p0+1-m)][Nlat], then zeros
  double *St = lds;                 // [cs][Nc], over Kv and Ts at a rotation's end
  const int cs = min(12, (R * L * kPtKs + R * 12 * FP) / Nc);
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  for (int i = tid; i < nmlt + (LMAX - 1) * Nlat; i += blockDim.x)
    Ml[i] = i < nmlt ? mlt[i] : 0.0;
  for (int i = tid; i < Nlat; i += blockDim.x) Qw[i] = qw[i];
  // kernel-value role: point (krow, lon) of the pass, pole t
  const bool kvt = tid < R * 2 * L;
  const int krow = kvt ? tid / (2 * L) : 0, lon = kvt ? (tid >> 1) % L : 0, t = tid & 1;
  // Fourier role (wave-uniform): row frow, output columns 16 tile .. 16 tile + 15
  const bool fw = wv < R * NT;
  const int frow = fw ? wv / NT : 0, tile = fw ? wv % NT : 0;
  const int li = lane & 15, lk = lane >> 4;
  // the wave's output column for lane li (-1: a padding column)
  int fcol = 16 * tile + li;
  if constexpr (FOLD) {
    fcol = -1;
    for (int cc = 0, seen = 0; cc < F; ++cc)
      if ((((cc + 1) / 2) & 1) == tile) {
        if (seen == li) fcol = cc;
        ++seen;
      }
  }
  const bool fcv = fw && fcol >= 0 && fcol < F;
  double mf[KS];
#pragma unroll
  for (int s = 0; s < KS; ++s) mf[s] = fcv ? mft[(long)fcol * L + 4 * s + lk] : 0.0;
  // Legendre role: component c, frequency row jj of the order-p0 layout,
  // whose block starts at off (after block 0 of length p0 + 1, the blocks
  // of order m' = 1 .. p0 come in pairs of length p0 + 1 - m')
  const bool lt = tid < 12 * F;
  const int c = lt ? tid / F : 0, jj = lt ? tid % F : 0;
  const int m = (jj + 1) / 2, len = p0 + 1 - m;
  const int mloff = Nlat * (m * (p0 + 1) - m * (m - 1) / 2);
  const int off = jj == 0 ? 0 : (p0 + 1) + 2 * ((m - 1) * (p0 + 1) - (m - 1) * m / 2) +
                                    ((jj % 2 == 0) ? (p0 + 1 - m) : 0);
  const int ng = (Nlat + R - 1) / R;
  const double eps = 2.220446049250313e-16;
  const double csl = 1.0 / (8.0 * kPi), cdl = 3.0 / (4.0 * kPi);
  // geometry and poles of the next (rotation, pass), loaded by every thread
  // at indices clamped into range: unconditional loads leave no value to
  // merge after a branch, so nothing waits for them before their use in the
  // next pass
  double g[9], tg[3];
  auto load = [&](long r_, int g_) {
    const long rc = r_ < nrot ? r_ : nrot - 1;
    const int lat_ = min(g_ * R + krow, Nlat - 1);
    const long b = rc * 3 * (long)M1 + (long)lat_ * L + lon;
    g[0] = X[b]; g[1] = X[b + M1]; g[2] = X[b + 2 * M1];
    g[3] = Xt[b]; g[4] = Xt[b + M1]; g[5] = Xt[b + 2 * M1];
    g[6] = Xp[b]; g[7] = Xp[b + M1]; g[8] = Xp[b + 2 * M1];
    for (int d = 0; d < 3; ++d) tg[d] = trg[(rc * 3 + d) * 2 + t];
  };
  load(blockIdx.x, 0);
  __syncthreads();  // Ml, Qw in
  for (long r = blockIdx.x; r < nrot; r += gridDim.x) {
    const double tx = tg[0], ty = tg[1], tz = tg[2];
    double acc[LMAX];
#pragma unroll
    for (int n = 0; n < LMAX; ++n) acc[n] = 0.0;
    for (int gi = 0; gi < ng; ++gi) {
      const double x0 = g[0], y0 = g[1], z0 = g[2];
      const double xt = g[3], yt = g[4], zt = g[5];
      const double xp = g[6], yp = g[7], zp = g[8];
      if (gi + 1 < ng)
        load(r, gi + 1);
      else
        load(r + gridDim.x, 0);
      const int lat = gi * R + krow;
      if (kvt && lat < Nlat) {
        const double nx = yt * zp - zt * yp, ny = zt * xp - xt * zp, nz = xt * yp - yt * xp;
        const double w = Qw[t ? lat : P1 - lat];
        const double dx = tx - x0, dy = ty - y0, dz = tz - z0;
        const double r2 = dx * dx + dy * dy + dz * dz;
        double rinv = 1.0 / sqrt(r2);
        if (r2 <= eps) rinv = 0;
        const double rinv2 = rinv * rinv;
        const double dd[6] = {dx * dx, dx * dy, dx * dz, dy * dy, dy * dz, dz * dz};
        double *kv = Kv + (krow * L + lon) * kPtKs + t * 6;
        if (layer == 1) {
          const double rn5 = cdl * w * (nx * dx + ny * dy + nz * dz) * rinv2 * rinv2 * rinv;
#pragma unroll
          for (int q = 0; q < 6; ++q) kv[q] = dd[q] * rn5;
        } else {
          const double area = sqrt(nx * nx + ny * ny + nz * nz);
          const double ar = csl * w * area * rinv;
          const double ar2 = ar * rinv2;
#pragma unroll
          for (int q = 0; q < 6; ++q)
            kv[q] = dd[q] * ar2 + ((q == 0 || q == 3 || q == 5) ? ar : 0.0);
        }
      }
      __syncthreads();  // Kv of the pass is in; the previous pass's T is consumed
      if (fw && gi * R + frow < Nlat) {
        const double *kv = Kv + frow * L * kPtKs + lk * kPtKs + li;
        v4d d = {0.0, 0.0, 0.0, 0.0}, e = {0.0, 0.0, 0.0, 0.0};
        double a[KS];
#pragma unroll
        for (int s = 0; s < KS; ++s) {
          double v = kv[4 * s * kPtKs];
          if constexpr (FOLD) {
            const double w = kv[(4 * s + L / 2) * kPtKs];
            v = tile ? v - w : v + w;
          }
          a[s] = li < 12 ? v : 0.0;
        }
#pragma unroll
        for (int s = 0; s < KS; s += 2) {  // two accumulation chains
          d = __builtin_amdgcn_mfma_f64_16x16x4f64(a[s], mf[s], d, 0, 0, 0);
          e = __builtin_amdgcn_mfma_f64_16x16x4f64(a[s + 1], mf[s + 1], e, 0, 0, 0);
        }
        d += e;
        if (fcol >= 0) {
#pragma unroll
          for (int i = 0; i < 3; ++i) Ts[(frow * 12 + lk + 4 * i) * FP + fcol] = d[i];
        }
      }
      __syncthreads();  // T of the pass is in
      if (lt) {
#pragma unroll 1
        for (int row = 0; row < R; ++row) {
          const int lat = gi * R + row;
          if (lat >= Nlat) break;
          // branch-free: degrees n >= len read finite values past the block
          // (the next blocks, then zeros) into accumulators that are never
          // written out
          const double f = Ts[(row * 12 + c) * FP + jj];
          const double *ml = Ml + mloff + lat;
#pragma unroll
          for (int n = 0; n < LMAX; ++n) acc[n] += f * ml[n * Nlat];
        }
      }
    }
    // the rotation's 12 output rows, cs rows per round through St
    double *out = K1 + r * 12 * (long)Nc;
    for (int c0 = 0; c0 < 12; c0 += cs) {
      __syncthreads();  // T, or the previous round of St, is read
      if (lt && c >= c0 && c < c0 + cs) {
#pragma unroll
        for (int n = 0; n < LMAX; ++n)
          if (n < len) St[(c - c0) * Nc + dpos[off + n]] = acc[n];
      }
      __syncthreads();
      const int nrow = min(cs, 12 - c0);
      for (int i = tid; i < nrow * Nc; i += blockDim.x) out[c0 * Nc + i] = St[i];
    }
    __syncthreads();  // St read before the next rotation's kernel values
  }
}

// M1arr[v][a*Ncoef+k][b][lat][lon] from K2[v][j1][j2][t2*6+c][k]
// (src/SphericalHarmonics.cc:1090-1131): t2=0 pole -> grid (j1, j2),
// t2=1 pole -> grid (p0-j1, p0+j2); c = sym(a, b).  For fixed (v, a, b) this
// is a transpose between k (contiguous in K2) and the grid point g
// (contiguous in M1arr): a block moves a 32 (k) x 64 (g) tile through LDS so
// that both the reads and the writes are whole contiguous runs.
constexpr int kRaK = 32, kRaG = 64;
template <typename T>
__global__ __launch_bounds__(256) void k_singular_rearrange(const T *__restrict__ K2, long Nv,
                                                            int p0, T *__restrict__ M1) {
  __shared__ T tile[kRaG][kRaK + 1];
  const int Ncoef = p0 * (p0 + 2), Ngrid = 2 * p0 * (p0 + 1);
  const int tk = (Ncoef + kRaK - 1) / kRaK, tg = (Ngrid + kRaG - 1) / kRaG;
  long bid = blockIdx.x;
  const int gt = (int)(bid % tg);
  bid /= tg;
  const int kt = (int)(bid % tk);
  bid /= tk;
  const int ab = (int)(bid % 9);
  const long v = bid / 9;
  if (v >= Nv) return;
  const int a = ab / 3, b = ab % 3;
  const int lo = a < b ? a : b, hi = a < b ? b : a;
  const int c = lo == 0 ? hi : (lo == 1 ? 2 + hi : 5);
  const int k0 = kt * kRaK, g0 = gt * kRaG;
  const int t = threadIdx.x;
#pragma unroll
  for (int q = 0; q < (kRaK * kRaG) / 256; ++q) {
    const int kk = t % kRaK, gg = t / kRaK + q * (256 / kRaK);
    const int g = g0 + gg, k = k0 + kk;
    T val = T(0);
    if (g < Ngrid && k < Ncoef) {
      const int lat = g / (2 * p0), lon = g % (2 * p0);
      int j1, j2, t2;
      if (lon < p0) {
        j1 = lat; j2 = lon; t2 = 0;
      } else {
        j1 = p0 - lat; j2 = lon - p0; t2 = 1;
      }
      val = K2[((((v * (p0 + 1) + j1) * (long)p0 + j2) * 12 + t2 * 6 + c) * (long)Ncoef) + k];
    }
    tile[gg][kk] = val;
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < (kRaK * kRaG) / 256; ++q) {
    const int gg = t % kRaG, kk = t / kRaG + q * (256 / kRaG);
    const int g = g0 + gg, k = k0 + kk;
    if (g < Ngrid && k < Ncoef)
      M1[(((v * 3 + a) * (long)Ncoef + k) * 3 + b) * Ngrid + g] = tile[gg][kk];
  }
}

// The rearrangement above and Grid2SHC(p0 -> p0) of its rows in one pass
// (FP64, p0 <= 16): out[v][a Nc + k][b Nc + c] = Grid2SHC of the grid function
// M1(v, a, k, b)[lat][lon] that k_singular_rearrange writes (the same K2
// element per grid point), then the Fourier and Legendre stages of
// sht_grid2shc (src/SphericalHarmonics.cc:151-235).  M1(v, a, k, b) depends on
// (a, b) only through the symmetric tensor component c6(a, b), so a workgroup
// takes (v, c6) and kRgK consecutive k and writes the rows of (a, b) and
// (b, a).  In rounds of four latitudes, wave w forms
//   Fs[w][k][j] = sum_lon M1(k)[lat][lon] MatFourierInv[lon][j],  lat = l0 + w,
// on v_mfma_f64_16x16x4f64 (lane l: A(l % 16, l / 16), B(l / 16, l % 16),
// results at rows l / 16 + 4 i, column l % 16), its A operand read from K2
// (16 consecutive k of one grid point: a 128-byte run per 16 lanes, the next
// round's loaded during this one) and MatFourierInv in registers; after a
// barrier every thread adds the round's Legendre terms
//   Fs[u][k][j_c] MatLegendreInv[m_c][lat][n_c]
// to its outputs (k, c), k = k0 + wave + 4 r, c = lane + 64 q.  Neither the
// rearranged grid values nor the Fourier sums pass through HBM.
constexpr int kRgK = 16;
constexpr int kRgMl = 17 * 153;  // MatLegendreInv doubles at p0 = 16
__global__ __launch_bounds__(256, 3) void k_rearrange_g2s(const double *__restrict__ K2,
                                                          long Nv, int p0,
                                                          const double *__restrict__ mfinv,
                                                          const double *__restrict__ mlinv,
                                                          double *__restrict__ out) {
  typedef double v4d __attribute__((ext_vector_type(4)));
  __shared__ double Fs[4][kRgK][33];
  __shared__ double Ml[kRgMl];
  const int Nc = p0 * (p0 + 2), Nlat = p0 + 1, L = 2 * p0;
  const int nml = Nlat * (p0 + 1) * (p0 + 2) / 2;
  const int nkt = (Nc + kRgK - 1) / kRgK;
  long bid = blockIdx.x;
  const int kt = (int)(bid % nkt);
  bid /= nkt;
  const int c6 = (int)(bid % 6);  // xx, xy, xz, yy, yz, zz
  const long v = bid / 6;
  if (v >= Nv) return;
  const int a = c6 < 3 ? 0 : (c6 < 5 ? 1 : 2), b = c6 < 3 ? c6 : (c6 < 5 ? c6 - 2 : 2);
  const int k0 = kt * kRgK;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int li = lane & 15, lk = lane >> 4;
  for (int i = tid; i < nml; i += 256) Ml[i] = mlinv[i];
  // B operand: MatFourierInv[lon = 4 s + lk][j = 16 ct + li]
  double bf[2][8];
#pragma unroll
  for (int ct = 0; ct < 2; ++ct)
#pragma unroll
    for (int s = 0; s < 8; ++s) {
      const int lon = 4 * s + lk, j = 16 * ct + li;
      bf[ct][s] = (lon < L && j < L) ? mfinv[lon * L + j] : 0.0;
    }
  // outputs c = lane + 64 q: block column j, MatLegendreInv offset of (m, n)
  // at latitude 0, and the block's row length (degrees)
  int cj[5], cb[5], cl[5];
#pragma unroll
  for (int q = 0; q < 5; ++q) {
    const int c = lane + 64 * q;
    cj[q] = -1;
    cb[q] = 0;
    cl[q] = 0;
    for (int jj = 0, off = 0; jj < L && c < Nc; ++jj) {
      const int m = (jj + 1) / 2, len = p0 + 1 - m;
      if (c < off + len) {
        cj[q] = jj;
        cl[q] = len;
        cb[q] = Nlat * (m * (p0 + 1) - m * (m - 1) / 2) + (c - off);
        break;
      }
      off += len;
    }
  }
  double acc[4][5];
#pragma unroll
  for (int r = 0; r < 4; ++r)
#pragma unroll
    for (int q = 0; q < 5; ++q) acc[r][q] = 0.0;
  const double *kv2 = K2 + v * (long)(p0 + 1) * p0 * 12 * Nc;
  const int kk = k0 + li;
  const bool kvalid = kk < Nc;
  // A operand of latitude lat: M1(kk)[lat][4 s + lk]
  double xa[8];
  auto load = [&](int lat) {
#pragma unroll
    for (int s = 0; s < 8; ++s) {
      const int lon = 4 * s + lk;
      double x = 0.0;
      if (kvalid && lon < L && lat < Nlat) {
        const bool t2 = lon >= p0;
        const int j1 = t2 ? p0 - lat : lat, j2 = t2 ? lon - p0 : lon;
        x = kv2[(((long)j1 * p0 + j2) * 12 + (t2 ? 6 : 0) + c6) * Nc + kk];
      }
      xa[s] = x;
    }
  };
  load(w);
  __syncthreads();  // Ml in
  for (int l0 = 0; l0 < Nlat; l0 += 4) {
    if (l0 + w < Nlat) {
      v4d d0 = {0.0, 0.0, 0.0, 0.0}, d1 = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int s = 0; s < 8; ++s) {
        d0 = __builtin_amdgcn_mfma_f64_16x16x4f64(xa[s], bf[0][s], d0, 0, 0, 0);
        d1 = __builtin_amdgcn_mfma_f64_16x16x4f64(xa[s], bf[1][s], d1, 0, 0, 0);
      }
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        Fs[w][lk + 4 * i][li] = d0[i];
        Fs[w][lk + 4 * i][16 + li] = d1[i];
      }
    }
    load(l0 + 4 + w);
    __syncthreads();  // the round's Fourier sums are in
    const int nu = min(4, Nlat - l0);
    for (int u = 0; u < nu; ++u) {
      const int lat = l0 + u;
#pragma unroll
      for (int q = 0; q < 5; ++q) {
        if (cj[q] < 0) continue;
        const double ml = Ml[cb[q] + lat * cl[q]];
#pragma unroll
        for (int r = 0; r < 4; ++r) acc[r][q] += Fs[u][w + 4 * r][cj[q]] * ml;
      }
    }
    __syncthreads();  // Fs read before the next round writes it
  }
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int k = k0 + w + 4 * r;
    if (k >= Nc) continue;
    double *o = out + ((((v * 3 + a) * (long)Nc + k) * 3 + b) * Nc);
    double *o2 = out + ((((v * 3 + b) * (long)Nc + k) * 3 + a) * Nc);
#pragma unroll
    for (int q = 0; q < 5; ++q)
      if (cj[q] >= 0) {
        o[lane + 64 * q] = acc[r][q];
        if (a != b) o2[lane + 64 * q] = acc[r][q];
      }
  }
}

template <typename T>
__global__ void k_copy_cast_grid(const double *__restrict__ in, long n, T *__restrict__ out) {
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    out[i] = (T)in[i];
}

namespace {
std::map<std::pair<int, int>, double *> g_sw;  // singular weights per (device, p1)
const double *sing_weights_dev(int p1) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return nullptr;
  std::lock_guard<std::mutex> lk(g_mu);
  auto key = std::make_pair(dev, p1);
  auto it = g_sw.find(key);
  if (it != g_sw.end()) return it->second;
  std::vector<double> w;
  singular_weights(p1, w);
  double *d = nullptr;
  if (hipMalloc(&d, sizeof(double) * w.size()) != hipSuccess) return nullptr;
  if (hipMemcpy(d, w.data(), sizeof(double) * w.size(), hipMemcpyHostToDevice) != hipSuccess)
    return nullptr;
  g_sw[key] = d;
  return d;
}
}  // namespace

// Elements of scratch the build needs per vesicle of a chunk; the fused
// pole stage needs neither the kernel values K0 nor the transposed
// synthesis's scratch.
static long singular_elems_per_ves(int p0, int p1, bool fused) {
  const long Nc = sh_ncoef(p0), Ng = sh_ngrid(p0), M1 = sh_ngrid(p1);
  const long nr = (long)(p0 + 1) * p0;  // rotations per vesicle
  long e = 0;
  e += 3 * Ng;                          // X0 in T
  e += 3 * Nc;                          // S0
  e += nr * 3 * Nc;                     // S (rotated)
  e += 3 * nr * 3 * M1;                 // X, Xt, Xp
  e += nr * 3 * 2;                      // poles
  if (!fused) e += nr * 12 * M1;        // K0
  e += nr * 12 * Nc;                    // K1
  e += nr * 12 * Nc;                    // K2
  e += 3 * Nc * 3 * Ng;                 // M1 rearranged
  long w1 = shc2grid_work(nr * 3, p0, p1);
  if (!fused) w1 = std::max(w1, grid2shc_work(nr * 12, p1, p0));
  w1 = std::max(w1, grid2shc_work(9 * Nc, p0, p0));
  w1 = std::max(w1, 3 * nr * 12 * Nc);  // rotation scratch
  e += w1;
  return e;
}

// Launch shape of k_pole_transpose for (p0, p1) on the current device:
// R latitude rows per pass (3 where the LDS allows, else fewer), threads
// covering the three roles, and the persistent grid's workgroups per CU;
// threads == 0 where the fused stage does not apply.
struct PoleShape {
  int R = 0, threads = 0, blocks = 0;
  size_t lds = 0;
  bool small = false;  // the <9, 16> instance (p0 <= 8, p1 = 16)
};

PoleShape pole_fused_shape(int p0, int p1) {
  PoleShape sh;
  // the instances' shapes: p0 <= 8 with p1 = 16, p0 <= 16 with p1 = 32
  // (the build's upsampled order 2 p0)
  const bool small = p0 <= 8 && p1 == 16;
  if (!small && !(p0 <= 16 && p1 == 32)) return sh;
  // MFMA wave tiles per row: the two parities of m at p1 = 32 (the kernel's
  // FOLD), else 16 consecutive columns
  const int L = 2 * p1, F = 2 * p0, FP = 16 * ((F + 15) / 16), NT = small ? FP / 16 : 2;
  const long Nc = sh_ncoef(p0);
  int dev = 0, maxlds = 0, cus = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&maxlds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) !=
          hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) {
    (void)hipGetLastError();
    return sh;
  }
  const void *kern = small ? (const void *)k_pole_transpose<9, 16>
                           : (const void *)k_pole_transpose<17, 32>;
  for (int R = 3; R >= 1; --R) {
    const int threads =
        64 * std::max({(int)cdiv(2L * R * L, 64), R * NT, (int)cdiv(12L * F, 64)});
    const long stage = (long)R * L * kPtKs + (long)R * 12 * FP;
    const size_t lds = sizeof(double) * (size_t)pole_lds_doubles(R, p0, p1, small ? 9 : 17);
    if (threads > 384 || stage < Nc || lds > (size_t)maxlds) continue;
    if (lds > 64 * 1024 && hipFuncSetAttribute(kern, hipFuncAttributeMaxDynamicSharedMemorySize,
                                               (int)lds) != hipSuccess) {
      (void)hipGetLastError();
      continue;
    }
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, kern, threads, lds) != hipSuccess) {
      (void)hipGetLastError();
      continue;
    }
    if (nb < 1) continue;
    sh.R = R;
    sh.threads = threads;
    sh.blocks = nb * cus;
    sh.lds = lds;
    sh.small = small;
    return sh;
  }
  return sh;
}

template <typename T>
int singular_build(const double *X0, long Nv, int p0, int p1, T *SL, T *DL, hipStream_t st) {
  if (p0 < 1 || p1 < 1 || p0 > 128 || p1 > 255 || Nv < 0) return VES3D_INVALID_PARAMETER;
  if (Nv == 0 || (!SL && !DL)) return VES3D_SUCCESS;
  if (!X0) return VES3D_INVALID_PARAMETER;
  const long Nc = sh_ncoef(p0), Ng = sh_ngrid(p0), M1 = sh_ngrid(p1);
  const long nr = (long)(p0 + 1) * p0;
  const double *qw = sing_weights_dev(p1);
  if (!qw) return VES3D_MEMORY_ERROR;
  // the fused stage serves the FP64 build: with FP32 transforms the
  // separate kernels' FP32 GEMMs are faster than its FP64 arithmetic
  const PoleShape ps = std::is_same<T, double>::value && g_pole_fused.load(std::memory_order_relaxed)
                           ? pole_fused_shape(p0, p1)
                           : PoleShape();
  const bool fused = ps.threads > 0;
  const long per = singular_elems_per_ves(p0, p1, fused);
  const long budget = (6L << 30) / (long)sizeof(T);
  long chunk = std::max(1L, std::min(Nv, budget / per));
  T *ws = (T *)workspace(kWsSing0, sizeof(T) * per * chunk);
  if (!ws) return VES3D_MEMORY_ERROR;
  hipEvent_t done = nullptr;
  {
    int dev = 0;
    V3D_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(g_mu);
    BuildOrder &o = g_build_order[dev];
    if (!o.done) V3D_HIP(hipEventCreateWithFlags(&o.done, hipEventDisableTiming));
    if (o.recorded) V3D_HIP(hipStreamWaitEvent(st, o.done, 0));
    done = o.done;
    o.recorded = true;
  }
  for (long v0 = 0; v0 < Nv; v0 += chunk) {
    const long cv = std::min(chunk, Nv - v0);
    const long rot = cv * nr;
    T *p = ws;
    T *x0 = p; p += cv * 3 * Ng;
    T *S0 = p; p += cv * 3 * Nc;
    T *S = p; p += rot * 3 * Nc;
    T *X = p; p += rot * 3 * M1;
    T *Xt = p; p += rot * 3 * M1;
    T *Xp = p; p += rot * 3 * M1;
    T *P = p; p += rot * 3 * 2;
    T *K0 = p; if (!fused) p += rot * 12 * M1;
    T *K1 = p; p += rot * 12 * Nc;
    T *K2 = p; p += rot * 12 * Nc;
    T *M1a = p; p += cv * 3 * Nc * 3 * Ng;
    T *w = p;
    const long nx0 = cv * 3 * Ng;
    k_copy_cast_grid<T><<<grid_for(nx0), 256, 0, st>>>(X0 + v0 * 3 * Ng, nx0, x0);
    V3D_CHECK(launch_status());
    V3D_CHECK(sht_grid2shc<T>(x0, cv * 3, p0, p0, S0, w, st));
    V3D_CHECK(rotate_all<T>(S0, cv, p0, 3, S, w, st));
    V3D_CHECK(sht_shc2grid<T>(S, rot * 3, p0, p1, X, Xt, Xp, w, st));
    V3D_CHECK(sht_shc2pole<T>(S, rot * 3, p0, P, st));
    for (int layer = 0; layer < 2; ++layer) {
      T *dst = layer == 0 ? SL : DL;
      if (!dst) continue;
      if constexpr (std::is_same<T, double>::value) {
        if (fused) {
          const double *mft = sht_mft<double>(p1, p0);
          const double *mlt = sht_mlt<double>(p1, p0);
          const RotMatsDev *rm = rot_mats(p0);
          if (!mft || !mlt || !rm) return VES3D_MEMORY_ERROR;
          const unsigned nb = (unsigned)std::min<long>(rot, ps.blocks);
          if (ps.small)
            k_pole_transpose<9, 16><<<nb, ps.threads, ps.lds, st>>>(
                X, Xt, Xp, P, qw, rot, p0, ps.R, mft, mlt, layer, rm->iperm, K1);
          else
            k_pole_transpose<17, 32><<<nb, ps.threads, ps.lds, st>>>(
                X, Xt, Xp, P, qw, rot, p0, ps.R, mft, mlt, layer, rm->iperm, K1);
          V3D_CHECK(launch_status());
        }
      }
      if (!fused) {
        k_pole_kernel<T><<<grid_for(rot * M1), 256, 0, st>>>(
            X, Xt, Xp, P, qw, rot, p1, layer == 0 ? K0 : nullptr, layer == 1 ? K0 : nullptr);
        V3D_CHECK(launch_status());
        V3D_CHECK(sht_shc2grid_transpose<T>(K0, rot * 12, p1, p0, K1, w, st));
      }
      V3D_CHECK(rotate_transpose<T>(K1, cv, p0, 12, K2, w, st, fused));
      if constexpr (std::is_same<T, double>::value) {
        if (fused) {  // rearrangement and Grid2SHC in one pass
          const ShMatsDev *sm = sh_mats(p0, p0);
          if (!sm) return VES3D_MEMORY_ERROR;
          const long g_blocks = cv * 6L * cdiv(Nc, kRgK);
          if (g_blocks > 0x7fffffffL) return VES3D_INVALID_PARAMETER;
          k_rearrange_g2s<<<(unsigned)g_blocks, 256, 0, st>>>(
              K2, cv, p0, sm->d64 + sm->h->Mfinv, sm->d64 + sm->h->Mlinv[0],
              dst + v0 * 9 * Nc * Nc);
          V3D_CHECK(launch_status());
          continue;
        }
      }
      const long ra_blocks = cv * 9L * cdiv(Nc, kRaK) * cdiv(Ng, kRaG);
      if (ra_blocks > 0x7fffffffL) return VES3D_INVALID_PARAMETER;
      k_singular_rearrange<T><<<(unsigned)ra_blocks, 256, 0, st>>>(K2, cv, p0, M1a);
      V3D_CHECK(launch_status());
      V3D_CHECK(sht_grid2shc<T>(M1a, cv * 9 * Nc, p0, p0, dst + v0 * 9 * Nc * Nc, w, st));
    }
  }
  return hip_status(hipEventRecord(done, st));
}

// ------------------------------------------------------------- self apply
// V[v][c] = sum_r F[v][r] SL[v][r][c] (+ G[v][r] DL[v][r][c]), r, c < 3 Ncoef
// (src/StokesVelocity.cc:397-447).  One workgroup per (vesicle, 256 columns);
// matrix rows stream through coalesced, the density vector sits in LDS.
template <typename TM>
__global__ __launch_bounds__(256) void k_self_apply(const TM *__restrict__ SL,
                                                     const double *__restrict__ F,
                                                     const TM *__restrict__ DL,
                                                     const double *__restrict__ G, int n,
                                                     double *__restrict__ V) {
  extern __shared__ double sh[];
  const long v = blockIdx.y;
  const int c = blockIdx.x * 256 + threadIdx.x;
  double *f = sh, *g = sh + n;
  for (int i = threadIdx.x; i < n; i += 256) {
    f[i] = SL ? F[v * n + i] : 0.0;
    g[i] = DL ? G[v * n + i] : 0.0;
  }
  __syncthreads();
  if (c >= n) return;
  double acc0 = 0, acc1 = 0, acc2 = 0, acc3 = 0;
  if (SL) {
    const TM *M = SL + v * (long)n * n + c;
    int r = 0;
    for (; r + 3 < n; r += 4) {
      acc0 += f[r] * (double)M[(long)r * n];
      acc1 += f[r + 1] * (double)M[(long)(r + 1) * n];
      acc2 += f[r + 2] * (double)M[(long)(r + 2) * n];
      acc3 += f[r + 3] * (double)M[(long)(r + 3) * n];
    }
    for (; r < n; ++r) acc0 += f[r] * (double)M[(long)r * n];
  }
  if (DL) {
    const TM *M = DL + v * (long)n * n + c;
    int r = 0;
    for (; r + 3 < n; r += 4) {
      acc0 += g[r] * (double)M[(long)r * n];
      acc1 += g[r + 1] * (double)M[(long)(r + 1) * n];
      acc2 += g[r + 2] * (double)M[(long)(r + 2) * n];
      acc3 += g[r + 3] * (double)M[(long)(r + 3) * n];
    }
    for (; r < n; ++r) acc0 += g[r] * (double)M[(long)r * n];
  }
  V[v * n + c] = (acc0 + acc1) + (acc2 + acc3);
}

// As k_self_apply for the singular build's operators, whose 3 x 3 blocks of
// Ncoef x Ncoef satisfy B_ab = B_ba (block (a, b) holds the rows of the
// symmetric tensor component c(a, b), formed once and written to both):
// thread cc reads the six blocks a <= b at column cc once and forms
// V[b Nc + cc] = sum_a sum_k F[a Nc + k] B_ab[k][cc] for b = 0, 1, 2 --
// two thirds of the general kernel's matrix reads.
template <typename TM>
__global__ __launch_bounds__(512) void k_self_apply_sym(const TM *__restrict__ SL,
                                                         const double *__restrict__ F,
                                                         const TM *__restrict__ DL,
                                                         const double *__restrict__ G, int nc,
                                                         double *__restrict__ V) {
  extern __shared__ double sh[];
  const int n = 3 * nc;
  const long v = blockIdx.y;
  const int cc = blockIdx.x * blockDim.x + threadIdx.x;
  double *f = sh, *g = sh + n;
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    f[i] = SL ? F[v * n + i] : 0.0;
    g[i] = DL ? G[v * n + i] : 0.0;
  }
  __syncthreads();
  if (cc >= nc) return;
  double v0 = 0, v1 = 0, v2 = 0;
  auto run = [&](const TM *M, const double *x) {
    const TM *m = M + v * (long)n * n + cc;  // element (row, b nc + cc) at row n + b nc
#pragma unroll 2
    for (int k = 0; k < nc; ++k) {
      const TM *r0 = m + (long)k * n, *r1 = m + (long)(nc + k) * n,
               *r2 = m + (long)(2 * nc + k) * n;
      const double b00 = (double)r0[0], b01 = (double)r0[nc], b02 = (double)r0[2 * nc];
      const double b11 = (double)r1[nc], b12 = (double)r1[2 * nc], b22 = (double)r2[2 * nc];
      const double x0 = x[k], x1 = x[nc + k], x2 = x[2 * nc + k];
      v0 += x0 * b00 + x1 * b01 + x2 * b02;
      v1 += x0 * b01 + x1 * b11 + x2 * b12;
      v2 += x0 * b02 + x1 * b12 + x2 * b22;
    }
  };
  if (SL) run(SL, f);
  if (DL) run(DL, g);
  V[v * n + cc] = v0;
  V[v * n + nc + cc] = v1;
  V[v * n + 2 * nc + cc] = v2;
}

template <typename TM>
int self_apply(const TM *SL, const double *F, const TM *DL, const double *G, long Nv, int p,
               double *V, hipStream_t st, bool sym) {
  if (p < 1 || Nv < 0 || Nv > 65535) return VES3D_INVALID_PARAMETER;
  if (Nv == 0) return VES3D_SUCCESS;
  if (!V || (SL && !F) || (DL && !G)) return VES3D_INVALID_PARAMETER;
  const int n = 3 * sh_ncoef(p);
  if (!SL && !DL) return hip_status(hipMemsetAsync(V, 0, sizeof(double) * n * Nv, st));
  if (sym) {
    const int nc = sh_ncoef(p), bs = 64 * (int)cdiv(std::min(nc, 512), 64);
    dim3 g2(cdiv(nc, bs), (unsigned)Nv);
    k_self_apply_sym<TM><<<g2, bs, 2 * n * sizeof(double), st>>>(SL, F, DL, G, nc, V);
    return launch_status();
  }
  dim3 grid(cdiv(n, 256), (unsigned)Nv);
  k_self_apply<TM><<<grid, 256, 2 * n * sizeof(double), st>>>(SL, F, DL, G, n, V);
  return launch_status();
}

template int rotate_all<double>(const double *, long, int, int, double *, double *, hipStream_t);
template int rotate_transpose<double>(const double *, long, int, int, double *, double *,
                                       hipStream_t, bool);
template int singular_build<double>(const double *, long, int, int, double *, double *, hipStream_t);
template int singular_build<float>(const double *, long, int, int, float *, float *, hipStream_t);
template int self_apply<double>(const double *, const double *, const double *, const double *,
                                long, int, double *, hipStream_t, bool);
template int self_apply<float>(const float *, const double *, const float *, const double *,
                               long, int, double *, hipStream_t, bool);

}  // namespace v3d

using namespace v3d;

extern "C" {
int ves3d_hip_rotate_all_f64(const double *S, size_t N, int p, int dof, double *S_out,
                             ves3d_stream_t stream) {
  if (!S || !S_out || dof < 1 || p < 1) return VES3D_INVALID_PARAMETER;
  const long nc = sh_ncoef(p);
  double *w = (double *)workspace(kWsSing1, sizeof(double) * (long)N * p * dof * nc * (p + 2));
  if (!w) return VES3D_MEMORY_ERROR;
  return rotate_all<double>(S, (long)N, p, dof, S_out, w, (hipStream_t)stream);
}
int ves3d_hip_rotate_transpose_f64(const double *S_in, size_t N, int p, int dof, double *S_out,
                                   ves3d_stream_t stream) {
  if (!S_in || !S_out || dof < 1 || p < 1) return VES3D_INVALID_PARAMETER;
  const long nc = sh_ncoef(p);
  double *w = (double *)workspace(kWsSing1, sizeof(double) * 2L * N * (p + 1) * p * dof * nc);
  if (!w) return VES3D_MEMORY_ERROR;
  return rotate_transpose<double>(S_in, (long)N, p, dof, S_out, w, (hipStream_t)stream, false);
}
int ves3d_hip_pole_kernel_shape(int p0, int p1, int *rows_host, int *threads_host,
                                int *workgroups_host) {
  if (p0 < 1 || p1 < 1 || p0 > 128 || p1 > 255) return VES3D_INVALID_PARAMETER;
  const PoleShape ps =
      g_pole_fused.load(std::memory_order_relaxed) ? pole_fused_shape(p0, p1) : PoleShape();
  if (rows_host) *rows_host = ps.R;
  if (threads_host) *threads_host = ps.threads;
  if (workgroups_host) *workgroups_host = ps.blocks;
  return VES3D_SUCCESS;
}
int ves3d_hip_set_pole_kernel(int fused) {
  if (fused != 0 && fused != 1) return VES3D_INVALID_PARAMETER;
  g_pole_fused.store(fused);
  return VES3D_SUCCESS;
}
int ves3d_hip_singular_build_f64(const double *X0, size_t nves, int p0, int p1, double *SL,
                                 double *DL, ves3d_stream_t stream) {
  return singular_build<double>(X0, (long)nves, p0, p1, SL, DL, (hipStream_t)stream);
}
int ves3d_hip_singular_build_f32(const double *X0, size_t nves, int p0, int p1, float *SL,
                                 float *DL, ves3d_stream_t stream) {
  return singular_build<float>(X0, (long)nves, p0, p1, SL, DL, (hipStream_t)stream);
}
int ves3d_hip_self_apply_f64(const double *SL, const double *F, const double *DL,
                             const double *G, size_t nves, int p, double *V,
                             ves3d_stream_t stream) {
  return self_apply<double>(SL, F, DL, G, (long)nves, p, V, (hipStream_t)stream, false);
}
int ves3d_hip_self_apply_f32(const float *SL, const double *F, const float *DL,
                             const double *G, size_t nves, int p, double *V,
                             ves3d_stream_t stream) {
  return self_apply<float>(SL, F, DL, G, (long)nves, p, V, (hipStream_t)stream, false);
}
}

// near.hip -- near-singular correction (NearSingular<Real>,
// src/NearSingular.cc) on gfx950, FP64 throughout, for the step's surface
// targets (trg_is_surf) or targets off the surfaces, with single- and
// double-layer densities.
//
// Setup (once per step, positions fixed):
//   k_bounds     per-vesicle source centroid / radius, target radius
//   k_cand       vesicle pairs whose bounding spheres come within r_near
//   k_search     per candidate pair: closest source of v for every target of
//                tv (LDS-staged sources), flag targets closer than r_near
//   k_compact    ordered compaction into the pair list (v, then target)
//   k_project    quadratic patch, Gauss-Newton projection, normal, repulsion,
//                interpolation points and Lagrange weights per pair
// Apply (every Stokes evaluation):
//   k_direct     direct SL sum of v at the target and its 7 interpolation
//                points; one workgroup per tile of <= 32 pairs of one v
//   k_combine    patch value of the self velocity + Lagrange extrapolation,
//                minus the direct value at the target
//   k_scatter    per-target ordered sum into the AxisMajor velocity
// The reference's Morton tree / LET (:80-790) is replaced by the bounding-
// sphere culling of k_cand and a brute-force scan over the candidate pair's
// sources: same pair set, no tree, no sort.
#include <algorithm>
#include <cmath>
#include <new>
#include <vector>

#include "near.hpp"

namespace v3d {
namespace {

constexpr int kDeg = 8;            // INTERP_DEG
constexpr int kTile = 32;          // pairs per k_direct workgroup (8 points each)
constexpr double kEps = 1.1102230246251565e-16;     // 2^-53, src/NearSingular.cc:1592-1599
constexpr double kEpsSqrt = 1.0536712127723509e-08;  // sqrt(2^-53), :1379-1387

inline unsigned nblk(long n) {
  long b = (n + 255) / 256;
  return (unsigned)std::max(1L, std::min(b, 65536L));
}
#define NGL(i, n) \
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < (long)(n); i += (long)gridDim.x * blockDim.x)

__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
  return v;
}

// InterPoints (src/NearSingular.cc:1356-1360)
__host__ __device__ inline double inter_points(int i) { return i ? 1.0 + (i - 1.0) / (kDeg - 2) : 0.0; }

// ---------------------------------------------------------------- setup
__global__ __launch_bounds__(256) void k_bounds(const double *__restrict__ src,
                                                const double *__restrict__ trg, long M, long ng,
                                                double *cen, double *rs2, double *rt2) {
  const long v = blockIdx.x;
  const double *s = src + v * M * 3;
  const double *t = trg + v * ng * 3;
  __shared__ double sh[3][4];
  __shared__ double c[3];
  __shared__ double mx[2][4];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  double a0 = 0, a1 = 0, a2 = 0;
  for (long k = threadIdx.x; k < M; k += 256) {
    a0 += s[3 * k];
    a1 += s[3 * k + 1];
    a2 += s[3 * k + 2];
  }
  a0 = wave_sum(a0);
  a1 = wave_sum(a1);
  a2 = wave_sum(a2);
  if (lane == 0) {
    sh[0][w] = a0;
    sh[1][w] = a1;
    sh[2][w] = a2;
  }
  __syncthreads();
  if (threadIdx.x < 3) {
    const int d = threadIdx.x;
    c[d] = (((sh[d][0] + sh[d][1]) + sh[d][2]) + sh[d][3]) * (1.0 / (double)M);
    cen[v * 3 + d] = c[d];
  }
  __syncthreads();
  double ms = 0, mt = 0;
  for (long k = threadIdx.x; k < M; k += 256) {
    const double dx = s[3 * k] - c[0], dy = s[3 * k + 1] - c[1], dz = s[3 * k + 2] - c[2];
    ms = fmax(ms, dx * dx + dy * dy + dz * dz);
  }
  for (long k = threadIdx.x; k < ng; k += 256) {
    const double dx = t[3 * k] - c[0], dy = t[3 * k + 1] - c[1], dz = t[3 * k + 2] - c[2];
    mt = fmax(mt, dx * dx + dy * dy + dz * dz);
  }
  ms = wave_max(ms);
  mt = wave_max(mt);
  if (lane == 0) {
    mx[0][w] = ms;
    mx[1][w] = mt;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    rs2[v] = fmax(fmax(mx[0][0], mx[0][1]), fmax(mx[0][2], mx[0][3]));
    rt2[v] = fmax(fmax(mx[1][0], mx[1][1]), fmax(mx[1][2], mx[1][3]));
  }
}

// Targets off the surfaces, in groups of tpg: centroid tcen and squared
// radius rt2 of every group.
__global__ __launch_bounds__(256) void k_tbounds(const double *__restrict__ trg, long n_trg,
                                                 long tpg, double *tcen, double *rt2) {
  const long g = blockIdx.x;
  const long a = g * tpg, b = min(a + tpg, n_trg);
  __shared__ double sh[3][4];
  __shared__ double c[3];
  __shared__ double mx[4];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  double a0 = 0, a1 = 0, a2 = 0;
  for (long k = a + threadIdx.x; k < b; k += 256) {
    a0 += trg[3 * k];
    a1 += trg[3 * k + 1];
    a2 += trg[3 * k + 2];
  }
  a0 = wave_sum(a0);
  a1 = wave_sum(a1);
  a2 = wave_sum(a2);
  if (lane == 0) {
    sh[0][w] = a0;
    sh[1][w] = a1;
    sh[2][w] = a2;
  }
  __syncthreads();
  if (threadIdx.x < 3) {
    const int d = threadIdx.x;
    c[d] = (((sh[d][0] + sh[d][1]) + sh[d][2]) + sh[d][3]) / (double)(b - a);
    tcen[g * 3 + d] = c[d];
  }
  __syncthreads();
  double mt = 0;
  for (long k = a + threadIdx.x; k < b; k += 256) {
    const double dx = trg[3 * k] - c[0], dy = trg[3 * k + 1] - c[1], dz = trg[3 * k + 2] - c[2];
    mt = fmax(mt, dx * dx + dy * dy + dz * dz);
  }
  mt = wave_max(mt);
  if (lane == 0) mx[w] = mt;
  __syncthreads();
  if (threadIdx.x == 0) rt2[g] = fmax(fmax(mx[0], mx[1]), fmax(mx[2], mx[3]));
}

// flags[tg][v] = 1 when some target of target group tg may lie within
// r_near of a source of v (bounding spheres: targets within sqrt(rt2[tg]) of
// tcen[tg], sources within sqrt(rs2[v]) of cen[v]; conservative margin).
// surf: the groups are the vesicles' own surface points (tcen = cen), which
// never pair with their own vesicle.
__global__ void k_cand(const double *__restrict__ cen, const double *__restrict__ rs2,
                       const double *__restrict__ tcen, const double *__restrict__ rt2, long ntg,
                       long nv, double r_near, int surf, unsigned char *flags) {
  NGL(i, ntg * nv) {
    const long tv = i / nv, v = i % nv;
    unsigned char f = 0;
    if (!surf || tv != v) {
      const double dx = tcen[3 * tv] - cen[3 * v], dy = tcen[3 * tv + 1] - cen[3 * v + 1],
                   dz = tcen[3 * tv + 2] - cen[3 * v + 2];
      // targets of tv lie within sqrt(rt2[tv]) of c_tv and sources of v
      // within sqrt(rs2[v]) of c_v, so a pair closer than r_near needs
      // |c_tv - c_v| < sqrt(rt2[tv]) + sqrt(rs2[v]) + r_near
      const double reach = (sqrt(rt2[tv]) + sqrt(rs2[v]) + r_near) * (1.0 + 1e-9) + 1e-12;
      f = (dx * dx + dy * dy + dz * dz) <= reach * reach;
    }
    flags[i] = f;
  }
}

// One workgroup per candidate pair (v, tv): closest source of v for every
// target of tv that can be within r_near of v.
// Targets of group tv are tv * ng .. min(tv * ng + ng, n_trg) - 1.
__global__ __launch_bounds__(256) void k_search(const double *__restrict__ src,
                                                const double *__restrict__ trg,
                                                const int *__restrict__ cand, long M, long ng,
                                                long n_trg, const double *__restrict__ cen,
                                                const double *__restrict__ rs2, double r_near,
                                                int *near_k, int *cnt) {
#pragma clang fp contract(off)
  const long c = blockIdx.x;
  const long v = cand[2 * c], tv = cand[2 * c + 1];
  const double *s = src + v * M * 3;
  const double cx = cen[3 * v], cy = cen[3 * v + 1], cz = cen[3 * v + 2];
  const double reach = (sqrt(rs2[v]) + r_near) * (1.0 + 1e-9) + 1e-12;
  const double rn2 = r_near * r_near;
  __shared__ double ls[3 * 256];
  __shared__ int wc[4];
  int count = 0;
  for (long j0 = 0; j0 < ng; j0 += 256) {
    const long j = j0 + threadIdx.x;
    double tx = 0, ty = 0, tz = 0;
    bool act = false;
    if (j < ng && tv * ng + j < n_trg) {
      const double *tp = trg + (tv * ng + j) * 3;
      tx = tp[0];
      ty = tp[1];
      tz = tp[2];
      const double dx = tx - cx, dy = ty - cy, dz = tz - cz;
      act = (dx * dx + dy * dy + dz * dz) <= reach * reach;
    }
    double best = INFINITY;
    int bk = -1;
    if (__syncthreads_or(act)) {
      for (long k0 = 0; k0 < M; k0 += 256) {
        const long kk = k0 + threadIdx.x;
        if (kk < M) {
          ls[3 * threadIdx.x] = s[3 * kk];
          ls[3 * threadIdx.x + 1] = s[3 * kk + 1];
          ls[3 * threadIdx.x + 2] = s[3 * kk + 2];
        }
        __syncthreads();
        const int kn = (int)std::min<long>(256, M - k0);
        if (act) {
          for (int k = 0; k < kn; ++k) {
            const double dx = tx - ls[3 * k], dy = ty - ls[3 * k + 1], dz = tz - ls[3 * k + 2];
            const double r2 = dx * dx + dy * dy + dz * dz;
            if (r2 < best) {
              best = r2;
              bk = (int)(k0 + k);
            }
          }
        }
        __syncthreads();
      }
    }
    const bool nr = act && best < rn2;
    if (j < ng) near_k[c * ng + j] = nr ? bk : -1;
    count += nr ? 1 : 0;
  }
  const double cs = wave_sum((double)count);
  if ((threadIdx.x & 63) == 0) wc[threadIdx.x >> 6] = (int)cs;
  __syncthreads();
  if (threadIdx.x == 0) cnt[c] = wc[0] + wc[1] + wc[2] + wc[3];
}

// Ordered compaction of near_k rows: pairs[(off[c] + r)] = (target, v, k).
__global__ __launch_bounds__(256) void k_compact(const int *__restrict__ near_k,
                                                 const int *__restrict__ off,
                                                 const int *__restrict__ cand, long ng,
                                                 int *pairs) {
  const long c = blockIdx.x;
  const int v = cand[2 * c], tv = cand[2 * c + 1];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  __shared__ int ws[4];
  __shared__ int running;
  if (threadIdx.x == 0) running = off[c];
  __syncthreads();
  for (long j0 = 0; j0 < ng; j0 += 256) {
    const long j = j0 + threadIdx.x;
    const int k = j < ng ? near_k[c * ng + j] : -1;
    const bool f = k >= 0;
    const unsigned long long b = __ballot(f);
    const int pre = __popcll(b & ((1ull << lane) - 1ull));
    if (lane == 0) ws[w] = __popcll(b);
    __syncthreads();
    int base = running;
    for (int q = 0; q < w; ++q) base += ws[q];
    if (f) {
      const long idx = base + pre;
      pairs[3 * idx] = (int)(tv * ng + j);
      pairs[3 * idx + 1] = v;
      pairs[3 * idx + 2] = k;
    }
    __syncthreads();
    if (threadIdx.x == 0) running += ws[0] + ws[1] + ws[2] + ws[3];
    __syncthreads();
  }
}

// ------------------------------------------------------ quadratic patch
// c[a][b][d]: coefficient of y^a x^b of component d; x runs along a stencil
// row (longitude), y down the rows (latitude) (src/NearSingular.cc:1560-1588).
struct Patch {
  double c[3][3][3];
};

__device__ void patch_from_mesh(const double m[9][3], Patch &P) {
#pragma clang fp contract(off)
  double tmp[3][3][3];
  for (int j = 0; j < 3; ++j)
    for (int d = 0; d < 3; ++d) {
      tmp[j][0][d] = m[j * 3 + 1][d];
      tmp[j][1][d] = (m[j * 3 + 2][d] - m[j * 3 + 0][d]) * 0.5;
      tmp[j][2][d] = (m[j * 3 + 2][d] + m[j * 3 + 0][d]) * 0.5 - m[j * 3 + 1][d];
    }
  for (int b = 0; b < 3; ++b)
    for (int d = 0; d < 3; ++d) {
      P.c[0][b][d] = tmp[1][b][d];
      P.c[1][b][d] = (tmp[2][b][d] - tmp[0][b][d]) * 0.5;
      P.c[2][b][d] = (tmp[2][b][d] + tmp[0][b][d]) * 0.5 - tmp[1][b][d];
    }
}

__device__ void patch_eval(const Patch &P, double x, double y, double out[3]) {
#pragma clang fp contract(off)
  const double xs[3] = {1.0, x, x * x}, ys[3] = {1.0, y, y * y};
  out[0] = out[1] = out[2] = 0.0;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      for (int d = 0; d < 3; ++d) out[d] += P.c[j][i][d] * xs[i] * ys[j];
}

__device__ void patch_grad(const Patch &P, double x, double y, double gx[3], double gy[3]) {
#pragma clang fp contract(off)
  const double xs[3] = {1.0, x, x * x}, ys[3] = {1.0, y, y * y};
  const double dxs[3] = {0.0, 1.0, 2 * x}, dys[3] = {0.0, 1.0, 2 * y};
  for (int d = 0; d < 3; ++d) gx[d] = gy[d] = 0.0;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      for (int d = 0; d < 3; ++d) {
        gx[d] += P.c[j][i][d] * dxs[i] * ys[j];
        gy[d] += P.c[j][i][d] * xs[i] * dys[j];
      }
}

__device__ __forceinline__ double dot3(const double a[3], const double b[3]) {
#pragma clang fp contract(off)
  return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

// Point fetchers: index k of the [M] per-vesicle point set (poles first).
struct PointsPM {  // point-major [nv][M][3]
  const double *a;
  long M;
  __device__ double operator()(long v, long k, int d) const { return a[(v * M + k) * 3 + d]; }
};
struct PointsAM {  // AxisMajor grid [nv][3][M-2] + poles [nv][3][2]
  const double *up, *pole;
  long mu;
  __device__ double operator()(long v, long k, int d) const {
    return k < 2 ? pole[(v * 3 + d) * 2 + k] : up[(v * 3 + d) * mu + (k - 2)];
  }
};

// QuadraticPatch::patch_mesh (src/NearSingular.cc:1734-1870)
template <typename F>
__device__ void patch_mesh(const F &pt, long v, int k, int pu, const double *__restrict__ mint,
                           const double *__restrict__ mfix, double out[9][3]) {
#pragma clang fp contract(off)
  const int K = 2 * pu, nlat = pu + 1;
  if (k < 2) {
    const long ring = 2 + (k == 0 ? 0 : (long)(nlat - 1) * K);
    double val[8][3];
    for (int o = 0; o < 8; ++o)
      for (int d = 0; d < 3; ++d) {
        double s = 0.0;
        for (int q = 0; q < K; ++q) s += mint[o * K + q] * pt(v, ring + q, d);
        val[o][d] = s;
      }
    const int ord0[9] = {0, 7, 6, 1, -1, 5, 2, 3, 4};
    const int ord1[9] = {0, 1, 2, 7, -1, 3, 6, 5, 4};
    double raw[9][3];
    for (int s = 0; s < 9; ++s) {
      const int o = k == 0 ? ord0[s] : ord1[s];
      for (int d = 0; d < 3; ++d) raw[s][d] = o < 0 ? pt(v, k, d) : val[o][d];
    }
    for (int r = 0; r < 9; ++r)
      for (int d = 0; d < 3; ++d) {
        double s = 0.0;
        for (int q = 0; q < 9; ++q) s += mfix[r * 9 + q] * raw[q][d];
        out[r][d] = s;
      }
    return;
  }
  const int k0 = (k - 2) / K, k1 = (k - 2) % K;
  const int cols[3] = {(k1 + K - 1) % K, k1, (k1 + 1) % K};
  for (int r = 0; r < 3; ++r) {
    const int lat = k0 - 1 + r;
    for (int c = 0; c < 3; ++c) {
      const long idx = lat < 0 ? 0 : (lat > nlat - 1 ? 1 : 2 + cols[c] + (long)lat * K);
      for (int d = 0; d < 3; ++d) out[r * 3 + c][d] = pt(v, idx, d);
    }
  }
}

// QuadraticPatch::project (src/NearSingular.cc:1589-1712).  The reference
// loops until convergence; here both loops are capped (never reached for a
// patch of a resolved surface) so that every thread terminates.
__device__ bool patch_project(const Patch &P, const double t[3], double &x, double &y,
                              double sc[3]) {
#pragma clang fp contract(off)
  x = 0.0;
  y = 0.0;
  patch_eval(P, x, y, sc);
  for (int outer = 0; outer < 256; ++outer) {
    double dR[3] = {t[0] - sc[0], t[1] - sc[1], t[2] - sc[2]};
    const double dR2 = dot3(dR, dR);
    double gx[3], gy[3];
    patch_grad(P, x, y, gx, gy);
    double n[3] = {gx[1] * gy[2] - gx[2] * gy[1], gx[2] * gy[0] - gx[0] * gy[2],
                   gx[0] * gy[1] - gx[1] * gy[0]};
    const double nn = sqrt(dot3(n, n));
    n[0] /= nn;
    n[1] /= nn;
    n[2] /= nn;
    const double dRn = dot3(dR, n);
    dR[0] -= dRn * n[0];
    dR[1] -= dRn * n[1];
    dR[2] -= dRn * n[2];
    if (sqrt(dot3(dR, dR)) < 1e-8) break;
    const double dxdx = dot3(gx, gx), dydy = dot3(gy, gy), dxdy = dot3(gx, gy);
    const double dxdR = dot3(gx, dR), dydR = dot3(gy, dR);
    const double det = dxdx * dydy - dxdy * dxdy;
    double dx = (dydy * dxdR - dxdy * dydR) / det;
    double dy = (dxdx * dydR - dxdy * dxdR) / det;
    if ((x <= -1.2 && dx < 0.0) || (x >= 1.2 && dx > 0.0) || (y <= -1.2 && dy < 0.0) ||
        (y >= 1.2 && dy > 0.0))
      break;
    if (x > -1.2 && x + dx < -1.2) {
      const double s = (-1.2 - x) / dx;
      dx *= s;
      dy *= s;
    }
    if (x < 1.2 && x + dx > 1.2) {
      const double s = (1.2 - x) / dx;
      dx *= s;
      dy *= s;
    }
    if (y > -1.2 && y + dy < -1.2) {
      const double s = (-1.2 - y) / dy;
      dx *= s;
      dy *= s;
    }
    if (y < 1.2 && y + dy > 1.2) {
      const double s = (1.2 - y) / dy;
      dx *= s;
      dy *= s;
    }
    for (int inner = 0; inner < 256; ++inner) {
      double e1[3], e2[3];
      patch_eval(P, x + dx * 0.5, y + dy * 0.5, e1);
      patch_eval(P, x + dx * 1.0, y + dy * 1.0, e2);
      const double r1[3] = {t[0] - e1[0], t[1] - e1[1], t[2] - e1[2]};
      const double r2[3] = {t[0] - e2[0], t[1] - e2[1], t[2] - e2[2]};
      const double d1 = dot3(r1, r1), d2 = dot3(r2, r2);
      const double c1 = -d2 + d1 * 4.0 - dR2 * 3.0;
      const double c2 = d2 * 2.0 - d1 * 4.0 + dR2 * 2.0;
      if (c2 > 0) {
        const double s = -0.5 * c1 / c2;
        if (s > 0.0 && s < 1.0) {
          dx *= s;
          dy *= s;
        }
      }
      patch_eval(P, x + dx, y + dy, sc);
      const double r[3] = {t[0] - sc[0], t[1] - sc[1], t[2] - sc[2]};
      if (dot3(r, r) < dR2) {
        x += dx;
        y += dy;
        break;
      }
      if (dx * dx * dxdx + dy * dy * dydy < 64 * kEps) break;
      dx *= 0.5;
      dy *= 0.5;
    }
    if (dx * dx * dxdx + dy * dy * dydy < 64 * kEps) break;
  }
  const double dR[3] = {t[0] - sc[0], t[1] - sc[1], t[2] - sc[2]};
  double gx[3], gy[3];
  patch_grad(P, x, y, gx, gy);
  double direc = 0;
  direc += dR[0] * gx[1] * gy[2];
  direc += dR[1] * gx[2] * gy[0];
  direc += dR[2] * gx[0] * gy[1];
  direc -= dR[0] * gx[2] * gy[1];
  direc -= dR[1] * gx[0] * gy[2];
  direc -= dR[2] * gx[1] * gy[0];
  return direc > 0;
}

// Per pair: projection, repulsion, interpolation points (index 0 = the
// target itself) and Lagrange weights (src/NearSingular.cc:1040-1098,
// :1426-1456, :1508-1540).  pxy = (x, y, 1 if the target sits on the patch,
// the factor of the double-layer density's jump in the patch value: +1/2,
// or -1/2 for an interior target off the surfaces, :1458-1469).  surf: the
// targets are surface points, whose interpolation always runs along the
// exterior normal.
__global__ void k_project(const double *__restrict__ src, const double *__restrict__ trg,
                          const int *__restrict__ pairs, long npairs, long M, int pu,
                          const double *__restrict__ mint, const double *__restrict__ mfix,
                          double r_near, double r2repul_inv, int surf, double *ipts, double *wts,
                          double *pxy, double *rep) {
#pragma clang fp contract(off)
  NGL(i, npairs) {
    const long t = pairs[3 * i], v = pairs[3 * i + 1];
    const int k = pairs[3 * i + 2];
    const double tc[3] = {trg[3 * t], trg[3 * t + 1], trg[3 * t + 2]};
    double mesh[9][3];
    patch_mesh(PointsPM{src, M}, v, k, pu, mint, mfix, mesh);
    Patch P;
    patch_from_mesh(mesh, P);
    double x, y, sc[3];
    const bool extr = patch_project(P, tc, x, y, sc);
    double pr[3];
    patch_eval(P, x, y, pr);
    double gx[3], gy[3];
    patch_grad(P, x, y, gx, gy);
    double n[3] = {gx[1] * gy[2] - gx[2] * gy[1], gx[2] * gy[0] - gx[0] * gy[2],
                   gx[0] * gy[1] - gx[1] * gy[0]};
    const double nn = sqrt(dot3(n, n));
    n[0] /= nn;
    n[1] /= nn;
    n[2] /= nn;
    const double dR[3] = {tc[0] - pr[0], tc[1] - pr[1], tc[2] - pr[2]};
    const double r = dot3(dR, n);
    double f = 0.0;
    if (r2repul_inv > 0 && r < r_near) f = exp(-r * r * r2repul_inv) / r * (1.0 - r / r_near);
    for (int d = 0; d < 3; ++d) rep[3 * i + d] = n[d] * f;

    double dn = sqrt(dot3(dR, dR));
    if (!extr && surf) dn = -dn;
    double oo = 1.0 / dn;
    if (fabs(dn) < kEpsSqrt) {
      dn = 0.0;
      oo = 0.0;
    }
    const double ix = dn / r_near;
    double *ip = ipts + i * kDeg * 3;
    for (int d = 0; d < 3; ++d) ip[d] = tc[d];
    for (int l = 1; l < kDeg; ++l) {
      const double xl = inter_points(l);
      for (int d = 0; d < 3; ++d) ip[3 * l + d] = pr[d] + dR[d] * oo * r_near * xl;
    }
    double *w = wts + i * kDeg;
    for (int l = 0; l < kDeg; ++l) {
      double yl = 1.0;
      const double x0 = inter_points(l);
      for (int q = 0; q < kDeg; ++q)
        if (q != l) yl *= (ix - inter_points(q)) / (x0 - inter_points(q));
      w[l] = ix == 0.0 ? (l == 0 ? 1.0 : 0.0) : yl;
    }
    pxy[4 * i] = x;
    pxy[4 * i + 1] = y;
    pxy[4 * i + 2] = ix == 0.0 ? 1.0 : 0.0;
    pxy[4 * i + 3] = (!extr && !surf) ? -0.5 : 0.5;
  }
}

// ------------------------------------------------------------- apply
// tiles[3 i] = (v, first pair, pair count <= kTile).  Thread (pair, l)
// evaluates at point l of its pair the single layer of q,
// 1/(8 pi) sum_s q_s / r + (q_s . d) d / r^3, and the double layer of
// qd = [g (3), n (3)] per source, 3/(4 pi) sum_s (d . n_s)(d . g_s) d / r^5,
// d = x - y_s (src/PVFMMInterface.cc:161-238, :284-369).
template <bool SL, bool DL>
__global__ __launch_bounds__(256) void k_direct(const double *__restrict__ src,
                                                const double *__restrict__ q,
                                                const double *__restrict__ qd,
                                                const double *__restrict__ ipts,
                                                const int *__restrict__ tiles, long M,
                                                double *dvel) {
  const int v = tiles[3 * blockIdx.x], first = tiles[3 * blockIdx.x + 1],
            count = tiles[3 * blockIdx.x + 2];
  const int pl = threadIdx.x / kDeg, l = threadIdx.x % kDeg;
  const bool act = pl < count;
  const long pt = ((long)(first + pl) * kDeg + l) * 3;
  double tx = 0, ty = 0, tz = 0;
  if (act) {
    tx = ipts[pt];
    ty = ipts[pt + 1];
    tz = ipts[pt + 2];
  }
  const double *s = src + (long)v * M * 3;
  const double *g = SL ? q + (long)v * M * 3 : nullptr;
  const double *gd = DL ? qd + (long)v * M * 6 : nullptr;
  constexpr int R = 3 + (SL ? 3 : 0) + (DL ? 6 : 0);  // doubles per staged source
  __shared__ double ls[R * 128];
  double u0 = 0, u1 = 0, u2 = 0, w0 = 0, w1 = 0, w2 = 0;
  for (long k0 = 0; k0 < M; k0 += 128) {
    if (threadIdx.x < 128) {
      const long kk = k0 + threadIdx.x;
      if (kk < M) {
        double *o = ls + R * threadIdx.x;
        for (int d = 0; d < 3; ++d) o[d] = s[3 * kk + d];
        if (SL)
          for (int d = 0; d < 3; ++d) o[3 + d] = g[3 * kk + d];
        if (DL)
          for (int d = 0; d < 6; ++d) o[(SL ? 6 : 3) + d] = gd[6 * kk + d];
      }
    }
    __syncthreads();
    const int kn = (int)std::min<long>(128, M - k0);
    if (act) {
      for (int k = 0; k < kn; ++k) {
        const double *o = ls + R * k;
        const double dx = tx - o[0], dy = ty - o[1], dz = tz - o[2];
        const double r2 = dx * dx + dy * dy + dz * dz;
        if (r2 > 0.0) {
          const double ri = 1.0 / sqrt(r2);
          if (SL) {
            const double fx = o[3], fy = o[4], fz = o[5];
            const double fd = (fx * dx + fy * dy + fz * dz) * ri * ri * ri;
            u0 += fx * ri + fd * dx;
            u1 += fy * ri + fd * dy;
            u2 += fz * ri + fd * dz;
          }
          if (DL) {
            const double *h = o + (SL ? 6 : 3);
            const double ri2 = ri * ri;
            const double c = (dx * h[3] + dy * h[4] + dz * h[5]) * (dx * h[0] + dy * h[1] + dz * h[2]) *
                             ri2 * ri2 * ri;
            w0 += c * dx;
            w1 += c * dy;
            w2 += c * dz;
          }
        }
      }
    }
    __syncthreads();
  }
  if (act) {
    const double sc = 1.0 / (8.0 * kPi), dc = 3.0 / (4.0 * kPi);
    dvel[pt] = u0 * sc + w0 * dc;
    dvel[pt + 1] = u1 * sc + w1 * dc;
    dvel[pt + 2] = u2 * sc + w2 * dc;
  }
}

// corr[i] = sum_l w_l y_l - direct(target), y_0 = patch value of the self
// velocity at the projection, y_l = direct value at interpolation point l.
// With a double-layer density (dup / dpole: the density on the upsampled
// grid and at the poles) the patch value carries its jump, pxy[4 i + 3]
// times the density (src/NearSingular.cc:1458-1469).
__global__ void k_combine(const int *__restrict__ pairs, long npairs, int pu,
                          const double *__restrict__ mint, const double *__restrict__ mfix,
                          const double *__restrict__ vup, const double *__restrict__ vpole,
                          const double *__restrict__ dup, const double *__restrict__ dpole,
                          const double *__restrict__ pxy, const double *__restrict__ wts,
                          const double *__restrict__ dvel, double *corr) {
#pragma clang fp contract(off)
  const long mu = 2L * pu * (pu + 1);
  NGL(i, npairs) {
    const long v = pairs[3 * i + 1];
    const int k = pairs[3 * i + 2];
    double mesh[9][3];
    patch_mesh(PointsAM{vup, vpole, mu}, v, k, pu, mint, mfix, mesh);
    if (dup) {
      double md[9][3];
      patch_mesh(PointsAM{dup, dpole, mu}, v, k, pu, mint, mfix, md);
      const double sc = pxy[4 * i + 3];
      for (int r = 0; r < 9; ++r)
        for (int d = 0; d < 3; ++d) mesh[r][d] += sc * md[r][d];
    }
    Patch P;
    patch_from_mesh(mesh, P);
    double pv[3];
    patch_eval(P, pxy[4 * i], pxy[4 * i + 1], pv);
    const double *dv = dvel + i * kDeg * 3;
    const double *w = wts + i * kDeg;
    for (int d = 0; d < 3; ++d) {
      double y = pv[d];
      if (pxy[4 * i + 2] == 0.0) {
        double s = 0.0;
        for (int l = 1; l < kDeg; ++l) s += w[l] * dv[3 * l + d];
        y = w[0] * pv[d] + s;
      }
      corr[3 * i + d] = y - dv[d];
    }
  }
}

// out[tv][d][j] += sum of per_pair over the pairs of target tgt[i] = tv ng + j,
// in pair order (perm[tptr[i] .. tptr[i+1]) lists them); ng == 0: the
// targets are point-major, out[t][d].
__global__ void k_scatter(const int *__restrict__ tgt, const int *__restrict__ tptr,
                          const int *__restrict__ perm, long nt, long ng,
                          const double *__restrict__ per_pair, double *out) {
  NGL(i, nt) {
    const long t = tgt[i];
    double s0 = 0, s1 = 0, s2 = 0;
    for (int b = tptr[i]; b < tptr[i + 1]; ++b) {
      const long pidx = perm[b];
      s0 += per_pair[3 * pidx];
      s1 += per_pair[3 * pidx + 1];
      s2 += per_pair[3 * pidx + 2];
    }
    if (ng) {
      const long tv = t / ng, j = t % ng;
      out[(tv * 3 + 0) * ng + j] += s0;
      out[(tv * 3 + 1) * ng + j] += s1;
      out[(tv * 3 + 2) * ng + j] += s2;
    } else {
      out[3 * t] += s0;
      out[3 * t + 1] += s1;
      out[3 * t + 2] += s2;
    }
  }
}

// Host tables: trigonometric ring interpolation M_patch_interp [8][2 pu]
// (:1738-1755) and the pole re-sampling M0 M1^-1 [9][9] (:1846-1868).
std::vector<double> ring_interp(int pu) {
  const int K = 2 * pu;
  std::vector<double> m(8 * K, 0.0);
  for (int k0 = 0; k0 < 8; ++k0)
    for (int k1 = 0; k1 < K; ++k1)
      for (int k2 = 0; k2 <= pu; ++k2) {
        const double c = (k2 == 0 || k2 == pu) ? 1.0 : 2.0;
        const double a0 = 2.0 * kPi * k2 * (k0 / 8.0), a1 = 2.0 * kPi * k2 * ((double)k1 / K);
        m[k0 * K + k1] += c * sin(a0) * sin(a1) / K;
        m[k0 * K + k1] += c * cos(a0) * cos(a1) / K;
      }
  return m;
}

std::vector<double> pole_fix() {
  double M0[9][9], M1[9][18];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      const double x0 = i - 1.0, y0 = j - 1.0;
      double x1 = x0, y1 = y0;
      if (x1 != 0.0 && y1 != 0.0) {
        x1 /= std::sqrt(2.0);
        y1 /= std::sqrt(2.0);
      }
      for (int m = 0; m < 3; ++m)
        for (int n = 0; n < 3; ++n) {
          M0[i * 3 + j][m * 3 + n] = std::pow(x0, m) * std::pow(y0, n);
          M1[i * 3 + j][m * 3 + n] = std::pow(x1, m) * std::pow(y1, n);
        }
    }
  // M1 is invertible (its pseudo-inverse is the inverse): Gauss-Jordan with
  // partial pivoting on [M1 | I].
  for (int r = 0; r < 9; ++r)
    for (int c = 0; c < 9; ++c) M1[r][9 + c] = r == c ? 1.0 : 0.0;
  for (int c = 0; c < 9; ++c) {
    int piv = c;
    for (int r = c + 1; r < 9; ++r)
      if (std::fabs(M1[r][c]) > std::fabs(M1[piv][c])) piv = r;
    for (int q = 0; q < 18; ++q) std::swap(M1[c][q], M1[piv][q]);
    const double d = M1[c][c];
    for (int q = 0; q < 18; ++q) M1[c][q] /= d;
    for (int r = 0; r < 9; ++r)
      if (r != c) {
        const double f = M1[r][c];
        for (int q = 0; q < 18; ++q) M1[r][q] -= f * M1[c][q];
      }
  }
  std::vector<double> out(81, 0.0);
  for (int r = 0; r < 9; ++r)
    for (int c = 0; c < 9; ++c)
      for (int q = 0; q < 9; ++q) out[r * 9 + c] += M0[r][q] * M1[q][9 + c];
  return out;
}

}  // namespace

template <typename T>
int NearG::Buf<T>::reserve(size_t count) {
  if (count <= n && p) return VES3D_SUCCESS;
  if (p) (void)hipFree(p);
  p = nullptr;
  n = 0;
  const size_t want = std::max<size_t>(count + count / 4, 16);
  if (hipMalloc(&p, sizeof(T) * want) != hipSuccess) {
    p = nullptr;
    return VES3D_MEMORY_ERROR;
  }
  n = want;
  return VES3D_SUCCESS;
}

template <typename T>
NearG::Buf<T>::~Buf() {
  if (p) (void)hipFree(p);
}

NearG::~NearG() = default;

int NearG::init(int p_, int pu_, long nv_, double repul_dist, double r_near_override,
                bool trg_is_surf_, long trg_per_group) {
  if (p_ < 1 || pu_ < 1 || nv_ < 1 || (!trg_is_surf_ && trg_per_group < 1))
    return VES3D_INVALID_PARAMETER;
  p = p_;
  pu = pu_;
  nv = nv_;
  trg_is_surf = trg_is_surf_;
  ng = 2L * p * (p + 1);
  tpg = trg_is_surf ? ng : trg_per_group;
  M = 2 + 2L * pu * (pu + 1);
  repul = repul_dist;
  r_near_fixed = r_near_override;
  V3D_CHECK(cen.reserve(3 * nv));
  V3D_CHECK(rs2.reserve(nv));
  V3D_CHECK(rt2.reserve(nv));
  V3D_CHECK(flags.reserve(nv * nv));
  const std::vector<double> mi = ring_interp(pu), mf = pole_fix();
  V3D_CHECK(mint.reserve(mi.size()));
  V3D_CHECK(mfix.reserve(mf.size()));
  V3D_HIP(hipMemcpy(mint.p, mi.data(), sizeof(double) * mi.size(), hipMemcpyHostToDevice));
  V3D_HIP(hipMemcpy(mfix.p, mf.data(), sizeof(double) * mf.size(), hipMemcpyHostToDevice));
  return VES3D_SUCCESS;
}

int NearG::setup(const double *src, const double *trg, hipStream_t st, long n_trg_) {
  // earlier applies may still read the per-pair buffers that reserve() below
  // can replace
  V3D_HIP(hipStreamSynchronize(st));
  npairs = ntrg_near = ntiles = 0;
  src_ = src;
  n_trg = trg_is_surf ? nv * ng : n_trg_;
  if (n_trg <= 0) return trg_is_surf ? VES3D_INVALID_PARAMETER : VES3D_SUCCESS;
  const long ntg = trg_is_surf ? nv : (n_trg + tpg - 1) / tpg;  // target groups
  V3D_CHECK(rt2.reserve(std::max(ntg, nv)));  // k_bounds writes nv entries first
  V3D_CHECK(flags.reserve(ntg * nv));
  k_bounds<<<(unsigned)nv, 256, 0, st>>>(src, trg_is_surf ? trg : src, M, trg_is_surf ? ng : 0,
                                          cen.p, rs2.p, rt2.p);
  V3D_CHECK(launch_status());
  const double *tc = cen.p;
  if (!trg_is_surf) {
    V3D_CHECK(tcen.reserve(3 * ntg));
    k_tbounds<<<(unsigned)ntg, 256, 0, st>>>(trg, n_trg, tpg, tcen.p, rt2.p);
    V3D_CHECK(launch_status());
    tc = tcen.p;
  }
  std::vector<double> h_rs2(nv);
  V3D_HIP(hipMemcpyAsync(h_rs2.data(), rs2.p, sizeof(double) * nv, hipMemcpyDeviceToHost, st));
  V3D_HIP(hipStreamSynchronize(st));
  double m2 = 0.0;
  for (double r : h_rs2) m2 = std::max(m2, r);
  // src/NearSingular.cc:83, :118-160
  r_near = r_near_fixed > 0 ? r_near_fixed : std::sqrt(m2) * (2.0 / std::sqrt((double)pu));

  k_cand<<<nblk(ntg * nv), 256, 0, st>>>(cen.p, rs2.p, tc, rt2.p, ntg, nv, r_near,
                                          trg_is_surf ? 1 : 0, flags.p);
  V3D_CHECK(launch_status());
  std::vector<unsigned char> h_flags(ntg * nv);
  V3D_HIP(hipMemcpyAsync(h_flags.data(), flags.p, ntg * nv, hipMemcpyDeviceToHost, st));
  V3D_HIP(hipStreamSynchronize(st));
  h_cand.clear();
  for (long v = 0; v < nv; ++v)
    for (long tv = 0; tv < ntg; ++tv)
      if (h_flags[tv * nv + v]) {
        h_cand.push_back((int)v);
        h_cand.push_back((int)tv);
      }
  const long ncand = (long)h_cand.size() / 2;
  if (ncand == 0) return VES3D_SUCCESS;
  V3D_CHECK(cand.reserve(2 * ncand));
  V3D_CHECK(cnt.reserve(ncand));
  V3D_CHECK(off.reserve(ncand));
  V3D_CHECK(near_k.reserve(ncand * tpg));
  V3D_HIP(hipMemcpyAsync(cand.p, h_cand.data(), sizeof(int) * 2 * ncand, hipMemcpyHostToDevice, st));
  k_search<<<(unsigned)ncand, 256, 0, st>>>(src, trg, cand.p, M, tpg, n_trg, cen.p, rs2.p, r_near,
                                              near_k.p, cnt.p);
  V3D_CHECK(launch_status());
  h_cnt.resize(ncand);
  V3D_HIP(hipMemcpyAsync(h_cnt.data(), cnt.p, sizeof(int) * ncand, hipMemcpyDeviceToHost, st));
  V3D_HIP(hipStreamSynchronize(st));
  h_off.resize(ncand);
  long tot = 0;
  for (long c = 0; c < ncand; ++c) {
    h_off[c] = (int)tot;
    tot += h_cnt[c];
  }
  npairs = tot;
  if (npairs == 0) return VES3D_SUCCESS;
  V3D_HIP(hipMemcpyAsync(off.p, h_off.data(), sizeof(int) * ncand, hipMemcpyHostToDevice, st));
  V3D_CHECK(pairs.reserve(3 * npairs));
  k_compact<<<(unsigned)ncand, 256, 0, st>>>(near_k.p, off.p, cand.p, tpg, pairs.p);
  V3D_CHECK(launch_status());
  h_pairs.resize(3 * npairs);
  V3D_HIP(hipMemcpyAsync(h_pairs.data(), pairs.p, sizeof(int) * 3 * npairs,
                         hipMemcpyDeviceToHost, st));
  V3D_HIP(hipStreamSynchronize(st));

  // k_direct tiles: runs of one source vesicle (pairs are ordered by v)
  std::vector<int> h_tiles;
  for (long a = 0; a < npairs;) {
    long b = a;
    while (b < npairs && h_pairs[3 * b + 1] == h_pairs[3 * a + 1]) ++b;
    for (long s = a; s < b; s += kTile) {
      h_tiles.push_back(h_pairs[3 * a + 1]);
      h_tiles.push_back((int)s);
      h_tiles.push_back((int)std::min<long>(kTile, b - s));
    }
    a = b;
  }
  ntiles = (long)h_tiles.size() / 3;
  // target CSR: pairs of each near target in pair order
  std::vector<int> order(npairs);
  for (long i = 0; i < npairs; ++i) order[i] = (int)i;
  std::stable_sort(order.begin(), order.end(),
                   [&](int a, int b) { return h_pairs[3 * a] < h_pairs[3 * b]; });
  std::vector<int> h_tgt, h_tptr;
  for (long i = 0; i < npairs; ++i) {
    const int t = h_pairs[3 * order[i]];
    if (h_tgt.empty() || h_tgt.back() != t) {
      h_tgt.push_back(t);
      h_tptr.push_back((int)i);
    }
  }
  h_tptr.push_back((int)npairs);
  ntrg_near = (long)h_tgt.size();
  V3D_CHECK(tiles.reserve(h_tiles.size()));
  V3D_CHECK(tgt.reserve(h_tgt.size()));
  V3D_CHECK(tptr.reserve(h_tptr.size()));
  V3D_CHECK(perm.reserve(npairs));
  V3D_HIP(hipMemcpyAsync(tiles.p, h_tiles.data(), sizeof(int) * h_tiles.size(),
                         hipMemcpyHostToDevice, st));
  V3D_HIP(hipMemcpyAsync(tgt.p, h_tgt.data(), sizeof(int) * h_tgt.size(), hipMemcpyHostToDevice, st));
  V3D_HIP(hipMemcpyAsync(tptr.p, h_tptr.data(), sizeof(int) * h_tptr.size(),
                         hipMemcpyHostToDevice, st));
  V3D_HIP(hipMemcpyAsync(perm.p, order.data(), sizeof(int) * npairs, hipMemcpyHostToDevice, st));

  V3D_CHECK(ipts.reserve(npairs * kDeg * 3));
  V3D_CHECK(dvel.reserve(npairs * kDeg * 3));
  V3D_CHECK(wts.reserve(npairs * kDeg));
  V3D_CHECK(pxy.reserve(npairs * 4));
  V3D_CHECK(rep.reserve(npairs * 3));
  V3D_CHECK(corr.reserve(npairs * 3));
  const double r2inv = repul > 0 ? 1.0 / (repul * repul) : 0.0;
  k_project<<<nblk(npairs), 256, 0, st>>>(src, trg, pairs.p, npairs, M, pu, mint.p, mfix.p, r_near,
                                           r2inv, trg_is_surf ? 1 : 0, ipts.p, wts.p, pxy.p,
                                           rep.p);
  V3D_CHECK(launch_status());
  // the host vectors above are released on return
  return hip_status(hipStreamSynchronize(st));
}

int NearG::add_repulsion(double *f, hipStream_t st) const {
  if (npairs == 0 || repul <= 0 || !trg_is_surf) return VES3D_SUCCESS;
  k_scatter<<<nblk(ntrg_near), 256, 0, st>>>(tgt.p, tptr.p, perm.p, ntrg_near, ng, rep.p, f);
  return launch_status();
}

int NearG::apply(const double *q, const double *qd, const double *vup, const double *vpole,
                 const double *dup, const double *dpole, double *vel, hipStream_t st) {
  if (npairs == 0) return VES3D_SUCCESS;
  if ((!q && !qd) || (qd && (!dup || !dpole))) return VES3D_INVALID_PARAMETER;
  if (q && qd)
    k_direct<true, true><<<(unsigned)ntiles, 256, 0, st>>>(src_, q, qd, ipts.p, tiles.p, M, dvel.p);
  else if (q)
    k_direct<true, false><<<(unsigned)ntiles, 256, 0, st>>>(src_, q, qd, ipts.p, tiles.p, M, dvel.p);
  else
    k_direct<false, true><<<(unsigned)ntiles, 256, 0, st>>>(src_, q, qd, ipts.p, tiles.p, M, dvel.p);
  V3D_CHECK(launch_status());
  k_combine<<<nblk(npairs), 256, 0, st>>>(pairs.p, npairs, pu, mint.p, mfix.p, vup, vpole,
                                           qd ? dup : nullptr, dpole, pxy.p, wts.p, dvel.p,
                                           corr.p);
  V3D_CHECK(launch_status());
  k_scatter<<<nblk(ntrg_near), 256, 0, st>>>(tgt.p, tptr.p, perm.p, ntrg_near,
                                              trg_is_surf ? ng : 0, corr.p, vel);
  return launch_status();
}

}  // namespace v3d

// ------------------------------------------------------------------ C ABI
using v3d::NearG;

extern "C" int ves3d_hip_near_create(int p, int p_up, size_t nves, double repul_dist,
                                     double r_near, void **handle_host) {
  if (!handle_host) return VES3D_INVALID_PARAMETER;
  *handle_host = nullptr;
  NearG *g = new (std::nothrow) NearG();
  if (!g) return VES3D_MEMORY_ERROR;
  const int s = g->init(p, p_up, (long)nves, repul_dist, r_near, true, 0);
  if (s != VES3D_SUCCESS) {
    delete g;
    return s;
  }
  *handle_host = g;
  return VES3D_SUCCESS;
}

extern "C" int ves3d_hip_near_destroy(void *handle) {
  delete static_cast<NearG *>(handle);
  return VES3D_SUCCESS;
}

extern "C" int ves3d_hip_near_set_coord(void *handle, const double *src, const double *trg,
                                        ves3d_stream_t stream) {
  if (!handle || !src || !trg) return VES3D_INVALID_PARAMETER;
  return static_cast<NearG *>(handle)->setup(src, trg, (hipStream_t)stream);
}

extern "C" int ves3d_hip_near_info(void *handle, size_t *npairs_host, double *r_near_host) {
  if (!handle) return VES3D_INVALID_PARAMETER;
  const NearG *g = static_cast<NearG *>(handle);
  if (npairs_host) *npairs_host = (size_t)g->npairs;
  if (r_near_host) *r_near_host = g->r_near;
  return VES3D_SUCCESS;
}

extern "C" int ves3d_hip_near_force_repul(void *handle, double *f, ves3d_stream_t stream) {
  if (!handle || !f) return VES3D_INVALID_PARAMETER;
  return static_cast<NearG *>(handle)->add_repulsion(f, (hipStream_t)stream);
}

extern "C" int ves3d_hip_near_apply(void *handle, const double *q, const double *vup,
                                    const double *vpole, double *vel, ves3d_stream_t stream) {
  if (!handle || !q || !vup || !vpole || !vel) return VES3D_INVALID_PARAMETER;
  return static_cast<NearG *>(handle)->apply(q, nullptr, vup, vpole, nullptr, nullptr, vel,
                                             (hipStream_t)stream);
}

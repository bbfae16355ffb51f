// stokes_direct.hip -- direct Stokes single/double-layer sums on gfx950.
//
// Two kernel families:
//
//  1. k_far_*: the all-pairs / far-field sum that stands in for the FMM
//     (StokesAlltoAll, src/CPUKernels.cc:966-1054, and the PVFMM pair kernels
//     src/PVFMMInterface.cc:161-238, :284-369).  One target per lane, 256
//     targets per workgroup; sources are staged through LDS in tiles of 256
//     and read back as wave-wide broadcasts.  The source range is cut into
//     `nsplit` slices; the slice index is the fastest-varying part of the
//     block index, so with nsplit a multiple of 8 every XCD (blocks are dealt
//     round-robin over the 8 XCDs) streams only its own slice of sources
//     through its L2.  Each slice writes an FP64 partial, and k_sum_slices adds
//     the partials in a fixed order (deterministic, no atomics).
//     The FP32 ("f32c") form stages each source tile relative to an FP64
//     anchor (the tile's first point) and shifts the lane's target by the same
//     anchor in FP64 before rounding, so coordinate differences keep FP32
//     relative accuracy independent of the box size; pair terms are summed in
//     FP32 over 64 sources and merged into an FP32 (hi, lo) TwoSum pair.
//
//  2. k_vesicle_*: the per-vesicle Device<>::DirectStokes /
//     DirectStokesDoubleLayer sums (src/CPUKernels.cc:8-118, :1059-1136) on
//     AxisMajor [ves][xyz][stride] data; one wave per 64 targets of one
//     vesicle, sources of that vesicle staged through LDS.
#include <atomic>
#include "common.hpp"
#include "implicit.hpp"
#include "runtime.hpp"

namespace v3d {

constexpr int FF_BLOCK = 256;
constexpr int FF_TILE = 256;
constexpr int FF_SUB = 64;  // FP32 form: TwoSum merge period (sources)

struct __align__(16) SrcRec64 {
  double x, y, z, _p0;
  double fx, fy, fz, _p1;
  double gx, gy, gz, _p2;
  double nx, ny, nz, _p3;
};
struct __align__(16) SrcRec32 {
  float x, y, z, _p0;
  float fx, fy, fz, _p1;
  float gx, gy, gz, _p2;
  float nx, ny, nz, _p3;
};

// Source range of one slice: whole groups when grouped, else an even split.
__device__ __forceinline__ void slice_range(long slice, int nsplit, long n_src,
                                            long spg, long n_groups, long &g0,
                                            long &g1) {
  if (spg > 0) {
    g0 = slice * n_groups / nsplit;
    g1 = (slice + 1) * n_groups / nsplit;
  } else {
    g0 = 0;
    g1 = 1;
  }
}

// 1/sqrt(x) for x > 0: the hardware estimate (v_rsq_f64, about 2^-20
// relative) refined by one third-order step y (1 + e/2 + 3e^2/8),
// e = 1 - x y^2, which leaves a rounding-level error.  x = 0 gives NaN here
// (callers select it away); no special-value fix-ups in the pair loop.
__device__ __forceinline__ double rsqrt_pair(double x) {
  const double y = __builtin_amdgcn_rsq(x);
  const double e = fma(y * -x, y, 1.0);
  return fma(y * e, fma(e, 0.375, 0.5), y);
}

// GUARD: drop pairs with r^2 <= eps2.  Without it a coincident pair gives
// NaN; the far field between vesicles runs without it (distinct vesicles'
// surfaces are disjoint, so every pair is at a positive distance).
template <bool SL, bool DL, bool GUARD>
__global__ __launch_bounds__(FF_BLOCK) void k_far_f64(
    const double *__restrict__ spos, const double *__restrict__ sld,
    const double *__restrict__ dld, long n_src, const double *__restrict__ tpos,
    long n_trg, long spg, long tpg, long n_groups, int nsplit, double eps2,
    double csl, double cdl, double *__restrict__ part) {
  __shared__ SrcRec64 sh[FF_TILE];
  const long slice = blockIdx.x % nsplit;
  const long tb = blockIdx.x / nsplit;
  const long t = tb * FF_BLOCK + threadIdx.x;
  const bool tvalid = t < n_trg;
  double tx = 0, ty = 0, tz = 0;
  if (tvalid) {
    tx = tpos[3 * t];
    ty = tpos[3 * t + 1];
    tz = tpos[3 * t + 2];
  }
  const bool grouped = spg > 0 && tpg > 0;
  const long my_g = (grouped && tvalid) ? t / tpg : -1;
  const long tfirst = tb * FF_BLOCK;
  const long tlast = min(tfirst + FF_BLOCK, n_trg) - 1;
  // the group holding all the block's targets (or -1), divided once
  const long own_b = (grouped && tfirst / tpg == tlast / tpg) ? tfirst / tpg : -1;

  double Sx = 0, Sy = 0, Sz = 0, Qx = 0, Qy = 0, Qz = 0;
  long g0, g1;
  slice_range(slice, nsplit, n_src, grouped ? spg : 0, n_groups, g0, g1);
  for (long g = g0; g < g1; ++g) {
    long sb, se;
    if (grouped) {
      if (g == own_b) continue;  // block-uniform skip
      sb = g * spg;
      se = sb + spg;
    } else {
      sb = slice * n_src / nsplit;
      se = (slice + 1) * n_src / nsplit;
    }
    const bool lane_on = tvalid && (my_g != g);
    for (long c0 = sb; c0 < se; c0 += FF_TILE) {
      const int cnt = (int)min((long)FF_TILE, se - c0);
      __syncthreads();
      for (int i = threadIdx.x; i < cnt; i += FF_BLOCK) {
        const long s = c0 + i;
        SrcRec64 r;
        r.x = spos[3 * s];
        r.y = spos[3 * s + 1];
        r.z = spos[3 * s + 2];
        if (SL) {
          r.fx = sld[3 * s];
          r.fy = sld[3 * s + 1];
          r.fz = sld[3 * s + 2];
        }
        if (DL) {
          r.gx = dld[6 * s];
          r.gy = dld[6 * s + 1];
          r.gz = dld[6 * s + 2];
          r.nx = dld[6 * s + 3];
          r.ny = dld[6 * s + 4];
          r.nz = dld[6 * s + 5];
        }
        sh[i] = r;
      }
      __syncthreads();
      if (lane_on) {
#pragma unroll 4
        for (int j = 0; j < cnt; ++j) {
          const double dx = tx - sh[j].x, dy = ty - sh[j].y, dz = tz - sh[j].z;
          const double r2 = dx * dx + dy * dy + dz * dz;
          // unconditional rsqrt + select: no per-pair branch in the loop
          const double rq = rsqrt_pair(r2);
          const double ri = GUARD ? ((r2 > eps2) ? rq : 0.0) : rq;
          const double ri2 = ri * ri;
          if (SL) {
            const double ip =
                (sh[j].fx * dx + sh[j].fy * dy + sh[j].fz * dz) * ri2;
            Sx += ri * (sh[j].fx + dx * ip);
            Sy += ri * (sh[j].fy + dy * ip);
            Sz += ri * (sh[j].fz + dz * ip);
          }
          if (DL) {
            const double p = (sh[j].nx * dx + sh[j].ny * dy + sh[j].nz * dz) *
                             (sh[j].gx * dx + sh[j].gy * dy + sh[j].gz * dz) *
                             (ri2 * ri2 * ri);
            Qx += dx * p;
            Qy += dy * p;
            Qz += dz * p;
          }
        }
      }
    }
  }
  if (tvalid) {
    double *o = part + (slice * n_trg + t) * 3;
    o[0] = csl * Sx + cdl * Qx;
    o[1] = csl * Sy + cdl * Qy;
    o[2] = csl * Sz + cdl * Qz;
  }
}

template <bool SL, bool DL, bool GUARD>
__global__ __launch_bounds__(FF_BLOCK) void k_far_f32c(
    const double *__restrict__ spos, const double *__restrict__ sld,
    const double *__restrict__ dld, long n_src, const double *__restrict__ tpos,
    long n_trg, long spg, long tpg, long n_groups, int nsplit, float eps2,
    double csl, double cdl, double *__restrict__ part) {
  __shared__ SrcRec32 sh[FF_TILE];
  const long slice = blockIdx.x % nsplit;
  const long tb = blockIdx.x / nsplit;
  const long t = tb * FF_BLOCK + threadIdx.x;
  const bool tvalid = t < n_trg;
  double tx = 0, ty = 0, tz = 0;
  if (tvalid) {
    tx = tpos[3 * t];
    ty = tpos[3 * t + 1];
    tz = tpos[3 * t + 2];
  }
  const bool grouped = spg > 0 && tpg > 0;
  const long my_g = (grouped && tvalid) ? t / tpg : -1;
  const long tfirst = tb * FF_BLOCK;
  const long tlast = min(tfirst + FF_BLOCK, n_trg) - 1;
  // the group holding all the block's targets (or -1), divided once
  const long own_b = (grouped && tfirst / tpg == tlast / tpg) ? tfirst / tpg : -1;

  // (hi, lo) compensated running sums
  float Sh[3] = {0, 0, 0}, Sl[3] = {0, 0, 0};
  float Qh[3] = {0, 0, 0}, Ql[3] = {0, 0, 0};
  long g0, g1;
  slice_range(slice, nsplit, n_src, grouped ? spg : 0, n_groups, g0, g1);
  for (long g = g0; g < g1; ++g) {
    long sb, se;
    if (grouped) {
      if (g == own_b) continue;
      sb = g * spg;
      se = sb + spg;
    } else {
      sb = slice * n_src / nsplit;
      se = (slice + 1) * n_src / nsplit;
    }
    const bool lane_on = tvalid && (my_g != g);
    for (long c0 = sb; c0 < se; c0 += FF_TILE) {
      const int cnt = (int)min((long)FF_TILE, se - c0);
      __syncthreads();
      const double ax = spos[3 * c0], ay = spos[3 * c0 + 1], az = spos[3 * c0 + 2];
      for (int i = threadIdx.x; i < cnt; i += FF_BLOCK) {
        const long s = c0 + i;
        SrcRec32 r;
        r.x = (float)(spos[3 * s] - ax);
        r.y = (float)(spos[3 * s + 1] - ay);
        r.z = (float)(spos[3 * s + 2] - az);
        if (SL) {
          r.fx = (float)sld[3 * s];
          r.fy = (float)sld[3 * s + 1];
          r.fz = (float)sld[3 * s + 2];
        }
        if (DL) {
          r.gx = (float)dld[6 * s];
          r.gy = (float)dld[6 * s + 1];
          r.gz = (float)dld[6 * s + 2];
          r.nx = (float)dld[6 * s + 3];
          r.ny = (float)dld[6 * s + 4];
          r.nz = (float)dld[6 * s + 5];
        }
        sh[i] = r;
      }
      __syncthreads();
      if (lane_on) {
        const float lx = (float)(tx - ax), ly = (float)(ty - ay), lz = (float)(tz - az);
        for (int j0 = 0; j0 < cnt; j0 += FF_SUB) {
          const int j1 = min(cnt, j0 + FF_SUB);
          float sx = 0, sy = 0, sz = 0, qx = 0, qy = 0, qz = 0;
#pragma unroll 4
          for (int j = j0; j < j1; ++j) {
            const float dx = lx - sh[j].x, dy = ly - sh[j].y, dz = lz - sh[j].z;
            const float r2 = dx * dx + dy * dy + dz * dz;
            const float rq = __frsqrt_rn(r2);
            const float ri = GUARD ? ((r2 > eps2) ? rq : 0.0f) : rq;
            const float ri2 = ri * ri;
            if (SL) {
              const float ip = (sh[j].fx * dx + sh[j].fy * dy + sh[j].fz * dz) * ri2;
              sx += ri * (sh[j].fx + dx * ip);
              sy += ri * (sh[j].fy + dy * ip);
              sz += ri * (sh[j].fz + dz * ip);
            }
            if (DL) {
              const float p = (sh[j].nx * dx + sh[j].ny * dy + sh[j].nz * dz) *
                              (sh[j].gx * dx + sh[j].gy * dy + sh[j].gz * dz) *
                              (ri2 * ri2 * ri);
              qx += dx * p;
              qy += dy * p;
              qz += dz * p;
            }
          }
          float s_, e_;
          if (SL) {
            two_sum(Sh[0], sx, s_, e_); Sh[0] = s_; Sl[0] += e_;
            two_sum(Sh[1], sy, s_, e_); Sh[1] = s_; Sl[1] += e_;
            two_sum(Sh[2], sz, s_, e_); Sh[2] = s_; Sl[2] += e_;
          }
          if (DL) {
            two_sum(Qh[0], qx, s_, e_); Qh[0] = s_; Ql[0] += e_;
            two_sum(Qh[1], qy, s_, e_); Qh[1] = s_; Ql[1] += e_;
            two_sum(Qh[2], qz, s_, e_); Qh[2] = s_; Ql[2] += e_;
          }
        }
      }
    }
  }
  if (tvalid) {
    double *o = part + (slice * n_trg + t) * 3;
#pragma unroll
    for (int c = 0; c < 3; ++c)
      o[c] = csl * ((double)Sh[c] + (double)Sl[c]) + cdl * ((double)Qh[c] + (double)Ql[c]);
  }
}

__global__ void k_sum_slices(const double *__restrict__ part, int nsplit, long n,
                             double *__restrict__ out) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double s = 0;
  for (int k = 0; k < nsplit; ++k) s += part[(long)k * n + i];
  out[i] = s;
}

// Number of source slices: a multiple of 8 (one per XCD), raised until the
// grid holds at least ~16 waves per SIMD of the 256-CU part.
static int choose_nsplit(long n_trg, long n_src, long n_groups) {
  const long tblocks = (n_trg + FF_BLOCK - 1) / FF_BLOCK;
  int ns = 8;
  while (ns < 64 && tblocks * ns * (FF_BLOCK / 64) < 16 * 1024) ns *= 2;
  long cap = n_groups > 0 ? n_groups : (n_src + FF_TILE - 1) / FF_TILE;
  if (cap < 1) cap = 1;
  if (ns > cap) ns = (int)cap;
  return ns;
}

template <bool F32, bool SL, bool DL>
static void launch_far(dim3 grid, hipStream_t st, const double *spos,
                       const double *sld, const double *dld, long n_src,
                       const double *tpos, long n_trg, long sp, long tp,
                       long n_groups, int ns, double eps2, double csl,
                       double cdl, double *part, bool guard) {
  if (F32) {
    if (guard)
      k_far_f32c<SL, DL, true><<<grid, FF_BLOCK, 0, st>>>(spos, sld, dld, n_src, tpos, n_trg, sp, tp,
                                                          n_groups, ns, (float)eps2, csl, cdl, part);
    else
      k_far_f32c<SL, DL, false><<<grid, FF_BLOCK, 0, st>>>(spos, sld, dld, n_src, tpos, n_trg, sp, tp,
                                                           n_groups, ns, (float)eps2, csl, cdl, part);
  } else {
    if (guard)
      k_far_f64<SL, DL, true><<<grid, FF_BLOCK, 0, st>>>(spos, sld, dld, n_src, tpos, n_trg, sp, tp,
                                                         n_groups, ns, eps2, csl, cdl, part);
    else
      k_far_f64<SL, DL, false><<<grid, FF_BLOCK, 0, st>>>(spos, sld, dld, n_src, tpos, n_trg, sp, tp,
                                                          n_groups, ns, eps2, csl, cdl, part);
  }
}

// Shared driver of the all-pairs and far-field entry points.
// part_fixed (optional): caller-provided scratch of part_cap bytes for the
// per-slice partial sums instead of the library's workspace; the slice count
// is then lowered until the partials fit.
template <bool F32>
static int far_eval(const double *spos, const double *sld, const double *dld,
                    size_t n_src, const double *tpos, double *tvel, size_t n_trg,
                    size_t spg, size_t tpg, double eps2, double csl, double cdl,
                    hipStream_t st, bool guard = true, double *part_fixed = nullptr,
                    size_t part_cap = 0) {
  if (n_trg == 0) return VES3D_SUCCESS;
  if (!tpos || !tvel) return VES3D_INVALID_PARAMETER;
  if (n_src == 0 || (!sld && !dld)) {
    V3D_HIP(hipMemsetAsync(tvel, 0, n_trg * 3 * sizeof(double), st));
    return VES3D_SUCCESS;
  }
  if (!spos) return VES3D_INVALID_PARAMETER;
  const bool grouped = spg > 0 && tpg > 0;
  long n_groups = 0;
  if (grouped) {
    if (n_src % spg || n_trg % tpg || n_src / spg != n_trg / tpg)
      return VES3D_INVALID_PARAMETER;
    n_groups = (long)(n_src / spg);
  }
  int ns = choose_nsplit((long)n_trg, (long)n_src, n_groups);
  double *part = nullptr;
  if (part_fixed) {
    const size_t per_slice = n_trg * 3 * sizeof(double);
    if (per_slice > part_cap) return VES3D_INVALID_PARAMETER;
    while ((size_t)ns * per_slice > part_cap) ns = ns > 8 ? ns / 2 : ns - 1;
    part = part_fixed;
  } else {
    part = (double *)workspace(kWsFarPartial, (size_t)ns * n_trg * 3 * sizeof(double));
  }
  if (!part) return VES3D_MEMORY_ERROR;
  const long tblocks = ((long)n_trg + FF_BLOCK - 1) / FF_BLOCK;
  const dim3 grid((unsigned)(tblocks * ns));
  const long sp = grouped ? (long)spg : 0, tp = grouped ? (long)tpg : 0;
  if (sld && dld)
    launch_far<F32, true, true>(grid, st, spos, sld, dld, n_src, tpos, n_trg, sp, tp, n_groups, ns, eps2, csl, cdl, part, guard);
  else if (sld)
    launch_far<F32, true, false>(grid, st, spos, sld, dld, n_src, tpos, n_trg, sp, tp, n_groups, ns, eps2, csl, cdl, part, guard);
  else
    launch_far<F32, false, true>(grid, st, spos, sld, dld, n_src, tpos, n_trg, sp, tp, n_groups, ns, eps2, csl, cdl, part, guard);
  V3D_CHECK(launch_status());
  const long n = (long)n_trg * 3;
  k_sum_slices<<<cdiv(n, 256), 256, 0, st>>>(part, ns, n, tvel);
  return launch_status();
}

// ---------------------------------------------------------------------------
// Far field between groups with the sources read through the scalar cache.
// Every lane of a wave uses the same source at the same time, so the source
// records are loaded with uniform addresses (s_load into SGPRs, which the
// VALU reads as operands): no LDS staging, no barriers, and the per-pair
// instruction stream is the arithmetic alone.  Records are packed once per
// call; the FP32 form stores each source relative to its group's first point
// (the group anchor, FP64), and the lane shifts its target by the same anchor
// in FP64 before rounding, as the LDS form does per tile.
struct __align__(16) FarRec64 {
  double x, y, z, fx, fy, fz, _p0, _p1;
};
struct __align__(16) FarRec32 {
  float x, y, z, fx, fy, fz, _p0, _p1;
};

__global__ void k_pack_far64(const double *__restrict__ spos, const double *__restrict__ sld,
                             long n, FarRec64 *__restrict__ rec) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  FarRec64 r;
  r.x = spos[3 * i];
  r.y = spos[3 * i + 1];
  r.z = spos[3 * i + 2];
  r.fx = sld[3 * i];
  r.fy = sld[3 * i + 1];
  r.fz = sld[3 * i + 2];
  r._p0 = r._p1 = 0.0;
  rec[i] = r;
}
// FP32 records relative to the group's anchor anc[3 (i / spg) apg] (the
// first fine source of the group, for both levels).
__global__ void k_pack_far32(const double *__restrict__ spos, const double *__restrict__ sld,
                             long n, long spg, const double *__restrict__ anc, long apg,
                             FarRec32 *__restrict__ rec) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double *a = anc + 3 * (i / spg) * apg;
  FarRec32 r;
  r.x = (float)(spos[3 * i] - a[0]);
  r.y = (float)(spos[3 * i + 1] - a[1]);
  r.z = (float)(spos[3 * i + 2] - a[2]);
  r.fx = (float)sld[3 * i];
  r.fy = (float)sld[3 * i + 1];
  r.fz = (float)sld[3 * i + 2];
  r._p0 = r._p1 = 0.0f;
  rec[i] = r;
}

// The r^2 bias that removes a target's own group from the FP64 sums (see
// kOwnGroupBias below): about 1e-15 |f| per own-group source.
constexpr double kOwnGroupBias64 = 1e30;

// Multi-level sources.  Every group (vesicle) carries its fine sources (spg
// records: the upsampled surface) and, optionally, coarser ones: level 1
// the surface at its own order with its own quadrature weights, level 2 a
// lower order still.  flags[w][g] names the level wave w (targets 64 w ..
// 64 w + 63) reads group g at: level l where every target of the wave is at
// least kappa_l times g's radius from g's centre, so that the coarser
// quadrature is as accurate as the fine one there.  The level, the record
// base and the count are wave-uniform (SGPRs), so the source loop keeps its
// scalar addressing.  A wave whose targets all belong to group g skips g.
struct WaveSpan {
  long row, first, last;  // flag row (clamped to the last wave), first and last target
  long own;               // the group that holds all the wave's targets, or -1
};
// The divisions run once per wave here, not once per group in the source
// loops (a 64-bit division is a long scalar sequence).
__device__ __forceinline__ WaveSpan wave_span(long tb, long n_trg, long tpg) {
  const long wv = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  const long w = tb * (FF_BLOCK / 64) + wv;
  const long n_waves = (n_trg + 63) / 64;
  WaveSpan r;
  r.row = w < n_waves ? w : n_waves - 1;
  r.first = w * 64;
  r.last = min(r.first + 64, n_trg) - 1;
  const long gf = r.first / tpg, gl = r.last / tpg;
  r.own = gf == gl ? gf : -1;
  return r;
}

// The source levels of the far field between groups: level 0 the fine
// records (spg per group), levels 1 and 2 coarser ones (cnt[l] per group,
// 0: absent); flags[w][g] names the level wave w reads group g at (null:
// level 0 everywhere).  The flags never name an absent level.
template <typename Rec>
struct FarLevels {
  const Rec *rec[3];
  long cnt[3];
  const unsigned char *flags;
  __device__ __forceinline__ int level(long row, long n_groups, long g) const {
    return flags ? (int)flags[row * n_groups + g] : 0;
  }
};

__global__ __launch_bounds__(FF_BLOCK) void k_far_groups_f64(
    const FarLevels<FarRec64> L, long spg, long n_groups, const double *__restrict__ tpos,
    long n_trg, long tpg, int nsplit, double csl, double *__restrict__ part) {
  const long slice = blockIdx.x % nsplit;
  const long tb = blockIdx.x / nsplit;
  const long t = tb * FF_BLOCK + threadIdx.x;
  const bool tvalid = t < n_trg;
  // lanes past the last target compute for it and store nothing, and a
  // target's own group enters through an r^2 bias, so the source loop's
  // control flow is uniform (its addressing stays scalar)
  const long tc = tvalid ? t : n_trg - 1;
  const double tx = tpos[3 * tc], ty = tpos[3 * tc + 1], tz = tpos[3 * tc + 2];
  const long my_g = tc / tpg;
  const WaveSpan ws = wave_span(tb, n_trg, tpg);
  double Sx = 0, Sy = 0, Sz = 0;
  long g0, g1;
  slice_range(slice, nsplit, 0, spg, n_groups, g0, g1);
  for (long g = g0; g < g1; ++g) {
    if (g == ws.own) continue;  // wave-uniform skip
    const double bias = my_g == g ? kOwnGroupBias64 : 0.0;
    const int lv = L.level(ws.row, n_groups, g);
    const FarRec64 *__restrict__ r = L.rec[lv] + g * L.cnt[lv];
    const int ns = (int)L.cnt[lv];
#pragma unroll 4
    for (int j = 0; j < ns; ++j) {
      const FarRec64 s = r[j];
      const double dx = tx - s.x, dy = ty - s.y, dz = tz - s.z;
      const double r2 = fma(dz, dz, fma(dy, dy, fma(dx, dx, bias)));
      const double ri = rsqrt_pair(r2);
      const double ri2 = ri * ri;
      const double ip = (s.fx * dx + s.fy * dy + s.fz * dz) * ri2;
      Sx += ri * (s.fx + dx * ip);
      Sy += ri * (s.fy + dy * ip);
      Sz += ri * (s.fz + dz * ip);
    }
  }
  if (tvalid) {
    double *o = part + (slice * n_trg + t) * 3;
    o[0] = csl * Sx;
    o[1] = csl * Sy;
    o[2] = csl * Sz;
  }
}

// FP32 partial sums merged into a TwoSum (hi, lo) pair, and the r^2 bias that
// removes a target's own group: each own-group source then adds about
// 1e-15 |f| (the coincident one included, without a NaN), far below FP32
// resolution of the sum.
struct Acc32 {
  float h[3], l[3];
};
__device__ __forceinline__ void acc_merge(Acc32 &a, float x, float y, float z) {
  float s_, e_;
  two_sum(a.h[0], x, s_, e_); a.h[0] = s_; a.l[0] += e_;
  two_sum(a.h[1], y, s_, e_); a.h[1] = s_; a.l[1] += e_;
  two_sum(a.h[2], z, s_, e_); a.h[2] = s_; a.l[2] += e_;
}
constexpr float kOwnGroupBias = 1e30f;

__global__ __launch_bounds__(FF_BLOCK) void k_far_groups_f32c(
    const FarLevels<FarRec32> L, const double *__restrict__ spos, long spg, long n_groups,
    const double *__restrict__ tpos, long n_trg, long tpg, int nsplit, double csl,
    double *__restrict__ part) {
  const long slice = blockIdx.x % nsplit;
  const long tb = blockIdx.x / nsplit;
  const long t = tb * FF_BLOCK + threadIdx.x;
  const bool tvalid = t < n_trg;
  // a lane past the last target computes for the last one and stores nothing,
  // so that control flow inside the group loop stays uniform
  const long tc = tvalid ? t : n_trg - 1;
  const double tx = tpos[3 * tc], ty = tpos[3 * tc + 1], tz = tpos[3 * tc + 2];
  const long my_g = tc / tpg;
  const WaveSpan ws = wave_span(tb, n_trg, tpg);
  Acc32 S = {{0, 0, 0}, {0, 0, 0}};
  long g0, g1;
  slice_range(slice, nsplit, 0, spg, n_groups, g0, g1);
  for (long g = g0; g < g1; ++g) {
    if (g == ws.own) continue;  // wave-uniform skip
    // a target's own group is not skipped lane by lane (that would make the
    // loop below divergent): its r^2 carries kOwnGroupBias instead
    const float bias = my_g == g ? kOwnGroupBias : 0.0f;
    const double *a = spos + 3 * g * spg;  // the group's anchor (every level)
    const float lx = (float)(tx - a[0]), ly = (float)(ty - a[1]), lz = (float)(tz - a[2]);
    const int lv = L.level(ws.row, n_groups, g);
    const FarRec32 *__restrict__ r = L.rec[lv] + g * L.cnt[lv];
    const long cnt = L.cnt[lv];
    auto chunk = [&](const FarRec32 *__restrict__ rr, int n, float &sx, float &sy,
                     float &sz) {
      auto pair = [&](const FarRec32 &src) {
        const float dx = lx - src.x, dy = ly - src.y, dz = lz - src.z;
        const float r2 = fmaf(dz, dz, fmaf(dy, dy, fmaf(dx, dx, bias)));
        const float ri = __frsqrt_rn(r2);
        const float ri2 = ri * ri;
        const float ip = (src.fx * dx + src.fy * dy + src.fz * dz) * ri2;
        sx += ri * (src.fx + dx * ip);
        sy += ri * (src.fy + dy * ip);
        sz += ri * (src.fz + dz * ip);
      };
      if (n == FF_SUB) {
#pragma unroll 8
        for (int j = 0; j < FF_SUB; ++j) pair(rr[j]);
      } else {
        for (int j = 0; j < n; ++j) pair(rr[j]);
      }
    };
    for (long j0 = 0; j0 < cnt; j0 += FF_SUB) {
      float sx = 0, sy = 0, sz = 0;
      chunk(r + j0, (int)min((long)FF_SUB, cnt - j0), sx, sy, sz);
      acc_merge(S, sx, sy, sz);
    }
  }
  if (tvalid) {
    double *o = part + (slice * n_trg + t) * 3;
#pragma unroll
    for (int c = 0; c < 3; ++c) o[c] = csl * ((double)S.h[c] + (double)S.l[c]);
  }
}

// Two targets per lane (t and t + FF_BLOCK): each source record loaded
// through the scalar cache serves two pairs.  A target must not see its own
// group's sources; instead of a lane-divergent skip (which would move the
// source addressing out of SGPRs), its r^2 carries a bias of 1e30 for that
// group, folded into the fma chain at no cost: each such source then adds
// about 1e-15 |f| (including the coincident one, without a NaN), some 1e-14
// of the sum and far below FP32 resolution.

__global__ __launch_bounds__(FF_BLOCK) void k_far_groups_f32c2(
    const FarRec32 *__restrict__ rec, const double *__restrict__ spos, long spg, long n_groups,
    const double *__restrict__ tpos, long n_trg, long tpg, int nsplit, double csl,
    double *__restrict__ part) {
  const long slice = blockIdx.x % nsplit;
  const long tb = blockIdx.x / nsplit;
  const long ta = tb * (2 * FF_BLOCK) + threadIdx.x, tbb = ta + FF_BLOCK;
  const bool va = ta < n_trg, vb = tbb < n_trg;
  const long ca = va ? ta : n_trg - 1, cb = vb ? tbb : n_trg - 1;
  const double Ax = tpos[3 * ca], Ay = tpos[3 * ca + 1], Az = tpos[3 * ca + 2];
  const double Bx = tpos[3 * cb], By = tpos[3 * cb + 1], Bz = tpos[3 * cb + 2];
  const long ga = ca / tpg, gbb = cb / tpg;
  const long tfirst = tb * (2 * FF_BLOCK);
  const long tlast = min(tfirst + 2 * FF_BLOCK, n_trg) - 1;
  const long own_b = tfirst / tpg == tlast / tpg ? tfirst / tpg : -1;
  Acc32 A = {{0, 0, 0}, {0, 0, 0}}, B = {{0, 0, 0}, {0, 0, 0}};
  long g0, g1;
  slice_range(slice, nsplit, 0, spg, n_groups, g0, g1);
  for (long g = g0; g < g1; ++g) {
    if (g == own_b) continue;  // block-uniform skip
    const double *a = spos + 3 * g * spg;
    const float ax = (float)(Ax - a[0]), ay = (float)(Ay - a[1]), az = (float)(Az - a[2]);
    const float bx = (float)(Bx - a[0]), by = (float)(By - a[1]), bz = (float)(Bz - a[2]);
    const FarRec32 *__restrict__ r = rec + g * spg;
    const float bias_a = ga == g ? kOwnGroupBias : 0.0f;
    const float bias_b = gbb == g ? kOwnGroupBias : 0.0f;
    for (long j0 = 0; j0 < spg; j0 += FF_SUB) {
      const long j1 = min(spg, j0 + FF_SUB);
      float ux = 0, uy = 0, uz = 0, vx = 0, vy = 0, vz = 0;
#pragma unroll 4
      for (long j = j0; j < j1; ++j) {
        const FarRec32 s = r[j];
        {
          const float dx = ax - s.x, dy = ay - s.y, dz = az - s.z;
          const float r2 = fmaf(dx, dx, fmaf(dy, dy, fmaf(dz, dz, bias_a)));
          const float ri = __frsqrt_rn(r2);
          const float ri2 = ri * ri;
          const float ip = (s.fx * dx + s.fy * dy + s.fz * dz) * ri2;
          ux += ri * (s.fx + dx * ip);
          uy += ri * (s.fy + dy * ip);
          uz += ri * (s.fz + dz * ip);
        }
        {
          const float dx = bx - s.x, dy = by - s.y, dz = bz - s.z;
          const float r2 = fmaf(dx, dx, fmaf(dy, dy, fmaf(dz, dz, bias_b)));
          const float ri = __frsqrt_rn(r2);
          const float ri2 = ri * ri;
          const float ip = (s.fx * dx + s.fy * dy + s.fz * dz) * ri2;
          vx += ri * (s.fx + dx * ip);
          vy += ri * (s.fy + dy * ip);
          vz += ri * (s.fz + dz * ip);
        }
      }
      acc_merge(A, ux, uy, uz);
      acc_merge(B, vx, vy, vz);
    }
  }
  if (va) {
    double *o = part + (slice * n_trg + ta) * 3;
#pragma unroll
    for (int c = 0; c < 3; ++c) o[c] = csl * ((double)A.h[c] + (double)A.l[c]);
  }
  if (vb) {
    double *o = part + (slice * n_trg + tbb) * 3;
#pragma unroll
    for (int c = 0; c < 3; ++c) o[c] = csl * ((double)B.h[c] + (double)B.l[c]);
  }
}

// Two densities on the same sources and targets in one pass (the step's
// right-hand side and the solver's first matvec): the pair geometry (d, r^2,
// 1/r, 1/r^2) is computed once and serves both Stokeslet sums.
struct __align__(16) FarRec32x2 {
  float x, y, z, ax, ay, az, bx, by, bz, _p0, _p1, _p2;
};
struct __align__(16) FarRec64x2 {
  double x, y, z, ax, ay, az, bx, by, bz, _p0, _p1, _p2;
};

__global__ void k_pack_far32x2(const double *__restrict__ spos, const double *__restrict__ s1,
                               const double *__restrict__ s2, long n, long spg,
                               const double *__restrict__ anc, long apg,
                               FarRec32x2 *__restrict__ rec) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double *a = anc + 3 * (i / spg) * apg;
  FarRec32x2 r;
  r.x = (float)(spos[3 * i] - a[0]);
  r.y = (float)(spos[3 * i + 1] - a[1]);
  r.z = (float)(spos[3 * i + 2] - a[2]);
  r.ax = (float)s1[3 * i];
  r.ay = (float)s1[3 * i + 1];
  r.az = (float)s1[3 * i + 2];
  r.bx = (float)s2[3 * i];
  r.by = (float)s2[3 * i + 1];
  r.bz = (float)s2[3 * i + 2];
  r._p0 = r._p1 = r._p2 = 0.0f;
  rec[i] = r;
}
__global__ void k_pack_far64x2(const double *__restrict__ spos, const double *__restrict__ s1,
                               const double *__restrict__ s2, long n,
                               FarRec64x2 *__restrict__ rec) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  FarRec64x2 r;
  r.x = spos[3 * i];
  r.y = spos[3 * i + 1];
  r.z = spos[3 * i + 2];
  r.ax = s1[3 * i];
  r.ay = s1[3 * i + 1];
  r.az = s1[3 * i + 2];
  r.bx = s2[3 * i];
  r.by = s2[3 * i + 1];
  r.bz = s2[3 * i + 2];
  r._p0 = r._p1 = r._p2 = 0.0;
  rec[i] = r;
}

__global__ __launch_bounds__(FF_BLOCK) void k_far_groups_f32c_x2(
    const FarLevels<FarRec32x2> L, const double *__restrict__ spos, long spg, long n_groups,
    const double *__restrict__ tpos, long n_trg, long tpg, int nsplit, double csl,
    double *__restrict__ part1, double *__restrict__ part2) {
  const long slice = blockIdx.x % nsplit;
  const long tb = blockIdx.x / nsplit;
  const long t = tb * FF_BLOCK + threadIdx.x;
  const bool tvalid = t < n_trg;
  const long tc = tvalid ? t : n_trg - 1;  // uniform control flow, as k_far_groups_f32c
  const double tx = tpos[3 * tc], ty = tpos[3 * tc + 1], tz = tpos[3 * tc + 2];
  const long my_g = tc / tpg;
  const WaveSpan ws = wave_span(tb, n_trg, tpg);
  Acc32 A = {{0, 0, 0}, {0, 0, 0}}, B = {{0, 0, 0}, {0, 0, 0}};
  long g0, g1;
  slice_range(slice, nsplit, 0, spg, n_groups, g0, g1);
  for (long g = g0; g < g1; ++g) {
    if (g == ws.own) continue;  // wave-uniform skip
    const float bias = my_g == g ? kOwnGroupBias : 0.0f;
    const double *a = spos + 3 * g * spg;
    const float lx = (float)(tx - a[0]), ly = (float)(ty - a[1]), lz = (float)(tz - a[2]);
    const int lv = L.level(ws.row, n_groups, g);
    const FarRec32x2 *__restrict__ r = L.rec[lv] + g * L.cnt[lv];
    const long n = L.cnt[lv];
    for (long j0 = 0; j0 < n; j0 += FF_SUB) {
      const int cnt = (int)min((long)FF_SUB, n - j0);
      const FarRec32x2 *__restrict__ rr = r + j0;
      float ux = 0, uy = 0, uz = 0, vx = 0, vy = 0, vz = 0;
      auto pair = [&](const FarRec32x2 &src) {
        const float dx = lx - src.x, dy = ly - src.y, dz = lz - src.z;
        const float r2 = fmaf(dz, dz, fmaf(dy, dy, fmaf(dx, dx, bias)));
        const float ri = __frsqrt_rn(r2);
        const float ri2 = ri * ri;
        const float ip = (src.ax * dx + src.ay * dy + src.az * dz) * ri2;
        ux += ri * (src.ax + dx * ip);
        uy += ri * (src.ay + dy * ip);
        uz += ri * (src.az + dz * ip);
        const float iq = (src.bx * dx + src.by * dy + src.bz * dz) * ri2;
        vx += ri * (src.bx + dx * iq);
        vy += ri * (src.by + dy * iq);
        vz += ri * (src.bz + dz * iq);
      };
      if (cnt == FF_SUB) {
#pragma unroll 4
        for (int j = 0; j < FF_SUB; ++j) pair(rr[j]);
      } else {
        for (int j = 0; j < cnt; ++j) pair(rr[j]);
      }
      acc_merge(A, ux, uy, uz);
      acc_merge(B, vx, vy, vz);
    }
  }
  if (tvalid) {
    double *o1 = part1 + (slice * n_trg + t) * 3, *o2 = part2 + (slice * n_trg + t) * 3;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      o1[c] = csl * ((double)A.h[c] + (double)A.l[c]);
      o2[c] = csl * ((double)B.h[c] + (double)B.l[c]);
    }
  }
}

__global__ __launch_bounds__(FF_BLOCK) void k_far_groups_f64_x2(
    const FarLevels<FarRec64x2> L, long spg, long n_groups, const double *__restrict__ tpos,
    long n_trg, long tpg, int nsplit, double csl, double *__restrict__ part1,
    double *__restrict__ part2) {
  const long slice = blockIdx.x % nsplit;
  const long tb = blockIdx.x / nsplit;
  const long t = tb * FF_BLOCK + threadIdx.x;
  const bool tvalid = t < n_trg;
  const long tc = tvalid ? t : n_trg - 1;  // uniform control flow, as k_far_groups_f64
  const double tx = tpos[3 * tc], ty = tpos[3 * tc + 1], tz = tpos[3 * tc + 2];
  const long my_g = tc / tpg;
  const WaveSpan ws = wave_span(tb, n_trg, tpg);
  double Sx = 0, Sy = 0, Sz = 0, Tx = 0, Ty = 0, Tz = 0;
  long g0, g1;
  slice_range(slice, nsplit, 0, spg, n_groups, g0, g1);
  for (long g = g0; g < g1; ++g) {
    if (g == ws.own) continue;  // wave-uniform skip
    const double bias = my_g == g ? kOwnGroupBias64 : 0.0;
    const int lv = L.level(ws.row, n_groups, g);
    const FarRec64x2 *__restrict__ r = L.rec[lv] + g * L.cnt[lv];
    const int ns = (int)L.cnt[lv];
#pragma unroll 2
    for (int j = 0; j < ns; ++j) {
      const FarRec64x2 s = r[j];
      const double dx = tx - s.x, dy = ty - s.y, dz = tz - s.z;
      const double r2 = fma(dz, dz, fma(dy, dy, fma(dx, dx, bias)));
      const double ri = rsqrt_pair(r2);
      const double ri2 = ri * ri;
      const double ip = (s.ax * dx + s.ay * dy + s.az * dz) * ri2;
      Sx += ri * (s.ax + dx * ip);
      Sy += ri * (s.ay + dy * ip);
      Sz += ri * (s.az + dz * ip);
      const double iq = (s.bx * dx + s.by * dy + s.bz * dz) * ri2;
      Tx += ri * (s.bx + dx * iq);
      Ty += ri * (s.by + dy * iq);
      Tz += ri * (s.bz + dz * iq);
    }
  }
  if (tvalid) {
    double *o1 = part1 + (slice * n_trg + t) * 3, *o2 = part2 + (slice * n_trg + t) * 3;
    o1[0] = csl * Sx;
    o1[1] = csl * Sy;
    o1[2] = csl * Sz;
    o2[0] = csl * Tx;
    o2[1] = csl * Ty;
    o2[2] = csl * Tz;
  }
}

// 0: LDS-staged source tiles (k_far_*), 1: scalar-cache source records
// (k_far_groups_*), 2: as 1 with two targets per lane in the FP32 form;
// selects the far field between groups.
static std::atomic<int> g_far_kind{1};

// ---------------------------------------------------------------------------
// Per-vesicle Device<>::DirectStokes / DirectStokesDoubleLayer.
constexpr int VS_BLOCK = 64;
constexpr int VS_TILE = 256;

// DLK=false: single layer, 1/(8pi) [w d / r + (w d . r) r / r^3]
// DLK=true : double layer, -3/(4pi) w (r.n)(r.d) r / r^5        (r = src - trg)
template <typename Real, bool DLK>
__global__ __launch_bounds__(VS_BLOCK) void k_vesicle_stokes(
    const Real *__restrict__ src, const Real *__restrict__ nrm,
    const Real *__restrict__ den, const Real *__restrict__ qw, long ss, long ts,
    long head, long tail, const Real *__restrict__ trg, Real *__restrict__ pot,
    Real scal) {
  __shared__ Real sh[VS_TILE][DLK ? 10 : 6];
  const long vt = blockIdx.y;
  const long t = head + (long)blockIdx.x * VS_BLOCK + threadIdx.x;
  const bool tvalid = t < tail;
  const Real *S = src + 3 * ss * vt;
  const Real *D = den + 3 * ss * vt;
  const Real *N = DLK ? nrm + 3 * ss * vt : nullptr;
  Real tx = 0, ty = 0, tz = 0;
  if (tvalid) {
    const Real *T = trg + 3 * ts * vt;
    tx = T[t];
    ty = T[ts + t];
    tz = T[2 * ts + t];
  }
  Real px = 0, py = 0, pz = 0;
  for (long c0 = 0; c0 < ss; c0 += VS_TILE) {
    const int cnt = (int)min((long)VS_TILE, ss - c0);
    __syncthreads();
    for (int i = threadIdx.x; i < cnt; i += VS_BLOCK) {
      const long s = c0 + i;
      const Real w = qw ? qw[s] : Real(1);
      sh[i][0] = S[s];
      sh[i][1] = S[ss + s];
      sh[i][2] = S[2 * ss + s];
      if (DLK) {
        sh[i][3] = D[s];
        sh[i][4] = D[ss + s];
        sh[i][5] = D[2 * ss + s];
        sh[i][6] = N[s];
        sh[i][7] = N[ss + s];
        sh[i][8] = N[2 * ss + s];
        sh[i][9] = w;
      } else {
        sh[i][3] = D[s] * w;
        sh[i][4] = D[ss + s] * w;
        sh[i][5] = D[2 * ss + s] * w;
      }
    }
    __syncthreads();
    if (tvalid) {
#pragma unroll 4
      for (int j = 0; j < cnt; ++j) {
        const Real dx = sh[j][0] - tx, dy = sh[j][1] - ty, dz = sh[j][2] - tz;
        const Real r2 = dx * dx + dy * dy + dz * dz;
        const Real ri = (r2 != Real(0)) ? Real(1) / sqrt(r2) : Real(0);
        if (DLK) {
          const Real ri2 = ri * ri;
          const Real rn = sh[j][6] * dx + sh[j][7] * dy + sh[j][8] * dz;
          const Real rf = sh[j][3] * dx + sh[j][4] * dy + sh[j][5] * dz;
          const Real c = rn * rf * (ri2 * ri2 * ri) * sh[j][9];
          px += dx * c;
          py += dy * c;
          pz += dz * c;
        } else {
          const Real cc = (dx * sh[j][3] + dy * sh[j][4] + dz * sh[j][5]) * ri * ri;
          px += (sh[j][3] + cc * dx) * ri;
          py += (sh[j][4] + cc * dy) * ri;
          pz += (sh[j][5] + cc * dz) * ri;
        }
      }
      Real *P = pot + 3 * ts * vt;
      P[t] = px * scal;
      P[ts + t] = py * scal;
      P[2 * ts + t] = pz * scal;
    }
  }
}

template <typename Real, bool DLK>
static int vesicle_stokes(const Real *src, const Real *nrm, const Real *den,
                          const Real *qw, size_t ss, size_t ts, size_t nv,
                          const Real *trg, size_t head, size_t tail, Real *pot,
                          hipStream_t st) {
  if (tail < head || tail > ts) return VES3D_INVALID_PARAMETER;
  if (nv == 0 || tail == head) return VES3D_SUCCESS;
  if (!src || !den || !trg || !pot || (DLK && !nrm)) return VES3D_INVALID_PARAMETER;
  if (nv > 65535) return VES3D_INVALID_PARAMETER;
  const Real scal = DLK ? Real(-3.0 / (4.0 * kPi)) : Real(1.0 / kPi / 8.0);
  dim3 grid(cdiv(tail - head, VS_BLOCK), (unsigned)nv);
  k_vesicle_stokes<Real, DLK><<<grid, VS_BLOCK, 0, st>>>(
      src, nrm, den, qw, (long)ss, (long)ts, (long)head, (long)tail, trg, pot, scal);
  return launch_status();
}

// Coincident-point guard of the PVFMMEval stand-in (r^2 <= 128 eps is skipped).
static const double kEvalEps2 = 128.0 * 2.220446049250313e-16;

// Bounding sphere of every group's fine sources: bnd[g] = (centre, radius),
// the centre the mean of the sources.  One workgroup per group.
__global__ __launch_bounds__(256) void k_group_bounds(const double *__restrict__ spos, long spg,
                                                      double4 *__restrict__ bnd) {
  const long g = blockIdx.x;
  const double *P = spos + 3 * g * spg;
  __shared__ double red[4][4];
  __shared__ double cen[3];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  double c[3] = {0, 0, 0};
  for (long j = threadIdx.x; j < spg; j += 256)
    for (int d = 0; d < 3; ++d) c[d] += P[3 * j + d];
  for (int d = 0; d < 3; ++d) {
    const double t = wave_sum(c[d]);
    if (lane == 0) red[wv][d] = t;
  }
  __syncthreads();
  if (threadIdx.x < 3)
    cen[threadIdx.x] =
        (((red[0][threadIdx.x] + red[1][threadIdx.x]) + red[2][threadIdx.x]) + red[3][threadIdx.x]) /
        (double)spg;
  __syncthreads();
  double m = 0;
  for (long j = threadIdx.x; j < spg; j += 256) {
    const double dx = P[3 * j] - cen[0], dy = P[3 * j + 1] - cen[1], dz = P[3 * j + 2] - cen[2];
    m = fmax(m, dx * dx + dy * dy + dz * dz);
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o, 64));
  if (lane == 0) red[wv][3] = m;
  __syncthreads();
  if (threadIdx.x == 0)
    bnd[g] = make_double4(cen[0], cen[1], cen[2],
                          sqrt(fmax(fmax(red[0][3], red[1][3]), fmax(red[2][3], red[3][3]))));
}

// flags[w][g] = 2 where every target of wave w is at least kappa2 R_g from
// c_g (kappa2 > 0), else 1 where at least kappa R_g, else 0.  One workgroup
// per wave: its 64 targets in LDS, the threads over the groups.
// counts[0] += the wave's level-1 (wave, group) pairs, counts[1] += its fine
// ones (groups the wave does not skip as its own), counts[2] += its level-2
// ones.
__global__ __launch_bounds__(256) void k_coarse_flags(const double *__restrict__ tpos, long n_trg,
                                                      long tpg, const double4 *__restrict__ bnd,
                                                      long n_groups, double kappa, double kappa2,
                                                      unsigned char *__restrict__ flags,
                                                      unsigned long long *__restrict__ counts) {
  const long w = blockIdx.x;
  __shared__ double T[64][3];
  __shared__ unsigned cnt[3];
  if (threadIdx.x < 64) {
    const long t = min(w * 64 + threadIdx.x, n_trg - 1);
    for (int d = 0; d < 3; ++d) T[threadIdx.x][d] = tpos[3 * t + d];
  }
  if (threadIdx.x < 3) cnt[threadIdx.x] = 0;
  __syncthreads();
  const long gf = (w * 64) / tpg, gl = min(w * 64 + 63, n_trg - 1) / tpg;
  unsigned nc = 0, nf = 0, nc2 = 0;
  for (long g = threadIdx.x; g < n_groups; g += 256) {
    const double4 b = bnd[g];
    double m = 1e300;
    for (int k = 0; k < 64; ++k) {
      const double dx = T[k][0] - b.x, dy = T[k][1] - b.y, dz = T[k][2] - b.z;
      m = fmin(m, dx * dx + dy * dy + dz * dz);
    }
    const double lim = kappa * b.w, lim2 = kappa2 * b.w;
    const int lv = (kappa2 > 0.0 && m >= lim2 * lim2) ? 2 : (m >= lim * lim ? 1 : 0);
    flags[w * n_groups + g] = (unsigned char)lv;
    if (lv == 2) ++nc2;
    else if (lv == 1) ++nc;
    else if (!(gf == g && gl == g)) ++nf;
  }
  atomicAdd(&cnt[0], nc);
  atomicAdd(&cnt[1], nf);
  atomicAdd(&cnt[2], nc2);
  __syncthreads();
  if (threadIdx.x < 3) atomicAdd(&counts[threadIdx.x], (unsigned long long)cnt[threadIdx.x]);
}

__global__ void k_zero_u64(unsigned long long *p, int n) {
  if ((int)threadIdx.x < n) p[threadIdx.x] = 0;
}

int far_field_sl_dl(bool f32c, const double *spos, const double *sld, const double *dld,
                    size_t n_src, const double *tpos, double *tvel, size_t n_trg, size_t spg,
                    size_t tpg, bool guard, hipStream_t st) {
  const double csl = 1.0 / (8.0 * kPi), cdl = 3.0 / (4.0 * kPi);
  const double eps2 = guard ? kEvalEps2 : 0.0;
  if (f32c)
    return far_eval<true>(spos, sld, dld, n_src, tpos, tvel, n_trg, spg, tpg, eps2, csl, cdl, st,
                          guard);
  return far_eval<false>(spos, sld, dld, n_src, tpos, tvel, n_trg, spg, tpg, eps2, csl, cdl, st,
                         guard);
}

int direct_eval_f64_fixed(const double *src_pos, const double *sl_den, const double *dl_den,
                          size_t n_src, const double *trg_pos, double *trg_vel, size_t n_trg,
                          size_t spg, size_t tpg, double *part, size_t part_bytes,
                          hipStream_t st) {
  if (!part) return VES3D_INVALID_PARAMETER;
  return far_eval<false>(src_pos, sl_den, dl_den, n_src, trg_pos, trg_vel, n_trg, spg, tpg,
                         kEvalEps2, 1.0 / (8.0 * kPi), 3.0 / (4.0 * kPi), st, true, part,
                         part_bytes);
}

// A level-2 set needs a level-1 set; two densities need both of each set.
static bool coarse_valid(const FarCoarse &c, bool two) {
  if (!c.pos || !c.den1 || !c.flags || c.cpg == 0 || (two && !c.den2)) return false;
  if (c.cpg2 == 0) return true;
  return c.pos2 && c.l2den1 && (!two || c.l2den2);
}

int far_coarse_flags(const double *spos, size_t n_src, size_t spg, const double *tpos,
                     size_t n_trg, size_t tpg, double kappa, double kappa2, unsigned char *flags,
                     unsigned long long *counts, hipStream_t st) {
  if (!spos || !tpos || !flags || !counts || spg == 0 || tpg == 0 || n_src % spg || n_trg == 0 ||
      !(kappa > 1.0) || !(kappa2 == 0.0 || kappa2 > kappa))
    return VES3D_INVALID_PARAMETER;
  const long n_groups = (long)(n_src / spg);
  double4 *bnd = (double4 *)workspace(kWsFarBounds, sizeof(double4) * n_groups);
  if (!bnd) return VES3D_MEMORY_ERROR;
  k_group_bounds<<<(unsigned)n_groups, 256, 0, st>>>(spos, (long)spg, bnd);
  V3D_CHECK(launch_status());
  k_zero_u64<<<1, 64, 0, st>>>(counts, 3);
  V3D_CHECK(launch_status());
  const long n_waves = ((long)n_trg + 63) / 64;
  k_coarse_flags<<<(unsigned)n_waves, 256, 0, st>>>(tpos, (long)n_trg, (long)tpg, bnd, n_groups,
                                                    kappa, kappa2, flags, counts);
  return launch_status();
}

int far_field_between_groups(bool f32c, const double *spos, const double *sld, size_t n_src,
                             const double *tpos, double *tvel, size_t n_trg, size_t spg,
                             size_t tpg, hipStream_t st, const FarCoarse *coarse) {
  if (spg == 0 || tpg == 0) return VES3D_INVALID_PARAMETER;
  const double csl = 1.0 / (8.0 * kPi), cdl = 3.0 / (4.0 * kPi);
  if (!coarse && g_far_kind.load() == 0) {
    if (f32c)
      return far_eval<true>(spos, sld, nullptr, n_src, tpos, tvel, n_trg, spg, tpg, 0.0, csl, cdl, st, false);
    return far_eval<false>(spos, sld, nullptr, n_src, tpos, tvel, n_trg, spg, tpg, 0.0, csl, cdl, st, false);
  }
  if (n_trg == 0) return VES3D_SUCCESS;
  if (!spos || !sld || !tpos || !tvel) return VES3D_INVALID_PARAMETER;
  if (n_src % spg || n_trg % tpg || n_src / spg != n_trg / tpg) return VES3D_INVALID_PARAMETER;
  if (coarse && !coarse_valid(*coarse, false)) return VES3D_INVALID_PARAMETER;
  const long n_groups = (long)(n_src / spg);
  const int ns = choose_nsplit((long)n_trg, (long)n_src, n_groups);
  const size_t rsz = f32c ? sizeof(FarRec32) : sizeof(FarRec64);
  double *part = (double *)workspace(kWsFarPartial, (size_t)ns * n_trg * 3 * sizeof(double));
  void *rec = workspace(kWsFarRecords, n_src * rsz);
  const long cpg = coarse ? (long)coarse->cpg : 0, n_c = cpg * n_groups;
  const long cpg2 = coarse ? (long)coarse->cpg2 : 0, n_c2 = cpg2 * n_groups;
  void *crec = coarse ? workspace(kWsFarCoarse, n_c * rsz) : nullptr;
  void *crec2 = cpg2 ? workspace(kWsFarCoarse2, n_c2 * rsz) : nullptr;
  if (!part || !rec || (coarse && !crec) || (cpg2 && !crec2)) return VES3D_MEMORY_ERROR;
  const unsigned char *cflag = coarse ? coarse->flags : nullptr;
  const long tblocks = ((long)n_trg + FF_BLOCK - 1) / FF_BLOCK;
  const dim3 grid((unsigned)(tblocks * ns));
  if (f32c) {
    k_pack_far32<<<cdiv((long)n_src, 256), 256, 0, st>>>(spos, sld, (long)n_src, (long)spg, spos,
                                                          (long)spg, (FarRec32 *)rec);
    V3D_CHECK(launch_status());
    if (coarse) {
      k_pack_far32<<<cdiv(n_c, 256), 256, 0, st>>>(coarse->pos, coarse->den1, n_c, cpg, spos,
                                                    (long)spg, (FarRec32 *)crec);
      V3D_CHECK(launch_status());
    }
    if (cpg2) {
      k_pack_far32<<<cdiv(n_c2, 256), 256, 0, st>>>(coarse->pos2, coarse->l2den1, n_c2, cpg2,
                                                     spos, (long)spg, (FarRec32 *)crec2);
      V3D_CHECK(launch_status());
    }
    if (!coarse && g_far_kind.load() == 2) {
      const long tb2 = ((long)n_trg + 2 * FF_BLOCK - 1) / (2 * FF_BLOCK);
      k_far_groups_f32c2<<<dim3((unsigned)(tb2 * ns)), FF_BLOCK, 0, st>>>(
          (const FarRec32 *)rec, spos, (long)spg, n_groups, tpos, (long)n_trg, (long)tpg, ns, csl,
          part);
    } else {
      const FarLevels<FarRec32> L{
          {(const FarRec32 *)rec, (const FarRec32 *)crec, (const FarRec32 *)crec2},
          {(long)spg, cpg, cpg2}, cflag};
      k_far_groups_f32c<<<grid, FF_BLOCK, 0, st>>>(L, spos, (long)spg, n_groups, tpos,
                                                   (long)n_trg, (long)tpg, ns, csl, part);
    }
  } else {
    k_pack_far64<<<cdiv((long)n_src, 256), 256, 0, st>>>(spos, sld, (long)n_src, (FarRec64 *)rec);
    V3D_CHECK(launch_status());
    if (coarse) {
      k_pack_far64<<<cdiv(n_c, 256), 256, 0, st>>>(coarse->pos, coarse->den1, n_c,
                                                    (FarRec64 *)crec);
      V3D_CHECK(launch_status());
    }
    if (cpg2) {
      k_pack_far64<<<cdiv(n_c2, 256), 256, 0, st>>>(coarse->pos2, coarse->l2den1, n_c2,
                                                     (FarRec64 *)crec2);
      V3D_CHECK(launch_status());
    }
    const FarLevels<FarRec64> L{
        {(const FarRec64 *)rec, (const FarRec64 *)crec, (const FarRec64 *)crec2},
        {(long)spg, cpg, cpg2}, cflag};
    k_far_groups_f64<<<grid, FF_BLOCK, 0, st>>>(L, (long)spg, n_groups, tpos, (long)n_trg,
                                                (long)tpg, ns, csl, part);
  }
  V3D_CHECK(launch_status());
  const long n = (long)n_trg * 3;
  k_sum_slices<<<cdiv(n, 256), 256, 0, st>>>(part, ns, n, tvel);
  return launch_status();
}

int far_field_between_groups2(bool f32c, const double *spos, const double *sld1,
                              const double *sld2, size_t n_src, const double *tpos, double *tvel1,
                              double *tvel2, size_t n_trg, size_t spg, size_t tpg,
                              hipStream_t st, const FarCoarse *coarse) {
  if (spg == 0 || tpg == 0) return VES3D_INVALID_PARAMETER;
  if (!coarse && g_far_kind.load() != 1) {  // the other kernels evaluate one density at a time
    V3D_CHECK(far_field_between_groups(f32c, spos, sld1, n_src, tpos, tvel1, n_trg, spg, tpg, st));
    return far_field_between_groups(f32c, spos, sld2, n_src, tpos, tvel2, n_trg, spg, tpg, st);
  }
  if (n_trg == 0) return VES3D_SUCCESS;
  if (!spos || !sld1 || !sld2 || !tpos || !tvel1 || !tvel2) return VES3D_INVALID_PARAMETER;
  if (n_src % spg || n_trg % tpg || n_src / spg != n_trg / tpg) return VES3D_INVALID_PARAMETER;
  if (coarse && !coarse_valid(*coarse, true)) return VES3D_INVALID_PARAMETER;
  const double csl = 1.0 / (8.0 * kPi);
  const long n_groups = (long)(n_src / spg);
  const int ns = choose_nsplit((long)n_trg, (long)n_src, n_groups);
  const size_t rsz = f32c ? sizeof(FarRec32x2) : sizeof(FarRec64x2);
  double *part = (double *)workspace(kWsFarPartial, (size_t)2 * ns * n_trg * 3 * sizeof(double));
  void *rec = workspace(kWsFarRecords, n_src * rsz);
  const long cpg = coarse ? (long)coarse->cpg : 0, n_c = cpg * n_groups;
  const long cpg2 = coarse ? (long)coarse->cpg2 : 0, n_c2 = cpg2 * n_groups;
  void *crec = coarse ? workspace(kWsFarCoarse, n_c * rsz) : nullptr;
  void *crec2 = cpg2 ? workspace(kWsFarCoarse2, n_c2 * rsz) : nullptr;
  if (!part || !rec || (coarse && !crec) || (cpg2 && !crec2)) return VES3D_MEMORY_ERROR;
  const unsigned char *cflag = coarse ? coarse->flags : nullptr;
  double *part2 = part + (size_t)ns * n_trg * 3;
  const long tblocks = ((long)n_trg + FF_BLOCK - 1) / FF_BLOCK;
  const dim3 grid((unsigned)(tblocks * ns));
  if (f32c) {
    k_pack_far32x2<<<cdiv((long)n_src, 256), 256, 0, st>>>(spos, sld1, sld2, (long)n_src,
                                                            (long)spg, spos, (long)spg,
                                                            (FarRec32x2 *)rec);
    V3D_CHECK(launch_status());
    if (coarse) {
      k_pack_far32x2<<<cdiv(n_c, 256), 256, 0, st>>>(coarse->pos, coarse->den1, coarse->den2,
                                                      n_c, cpg, spos, (long)spg,
                                                      (FarRec32x2 *)crec);
      V3D_CHECK(launch_status());
    }
    if (cpg2) {
      k_pack_far32x2<<<cdiv(n_c2, 256), 256, 0, st>>>(coarse->pos2, coarse->l2den1,
                                                       coarse->l2den2, n_c2, cpg2, spos,
                                                       (long)spg, (FarRec32x2 *)crec2);
      V3D_CHECK(launch_status());
    }
    const FarLevels<FarRec32x2> L{
        {(const FarRec32x2 *)rec, (const FarRec32x2 *)crec, (const FarRec32x2 *)crec2},
        {(long)spg, cpg, cpg2}, cflag};
    k_far_groups_f32c_x2<<<grid, FF_BLOCK, 0, st>>>(L, spos, (long)spg, n_groups, tpos,
                                                    (long)n_trg, (long)tpg, ns, csl, part, part2);
  } else {
    k_pack_far64x2<<<cdiv((long)n_src, 256), 256, 0, st>>>(spos, sld1, sld2, (long)n_src,
                                                            (FarRec64x2 *)rec);
    V3D_CHECK(launch_status());
    if (coarse) {
      k_pack_far64x2<<<cdiv(n_c, 256), 256, 0, st>>>(coarse->pos, coarse->den1, coarse->den2,
                                                      n_c, (FarRec64x2 *)crec);
      V3D_CHECK(launch_status());
    }
    if (cpg2) {
      k_pack_far64x2<<<cdiv(n_c2, 256), 256, 0, st>>>(coarse->pos2, coarse->l2den1,
                                                       coarse->l2den2, n_c2, (FarRec64x2 *)crec2);
      V3D_CHECK(launch_status());
    }
    const FarLevels<FarRec64x2> L{
        {(const FarRec64x2 *)rec, (const FarRec64x2 *)crec, (const FarRec64x2 *)crec2},
        {(long)spg, cpg, cpg2}, cflag};
    k_far_groups_f64_x2<<<grid, FF_BLOCK, 0, st>>>(L, (long)spg, n_groups, tpos, (long)n_trg,
                                                   (long)tpg, ns, csl, part, part2);
  }
  V3D_CHECK(launch_status());
  const long n = (long)n_trg * 3;
  k_sum_slices<<<cdiv(n, 256), 256, 0, st>>>(part, ns, n, tvel1);
  V3D_CHECK(launch_status());
  k_sum_slices<<<cdiv(n, 256), 256, 0, st>>>(part2, ns, n, tvel2);
  return launch_status();
}

}  // namespace v3d

// ---------------------------------------------------------------------------
// C ABI
using namespace v3d;

extern "C" {

int ves3d_hip_direct_stokes_f64(const double *src, const double *den,
                                const double *qw, size_t src_stride,
                                size_t trg_stride, size_t n_surfs,
                                const double *trg, size_t head, size_t tail,
                                double *pot, ves3d_stream_t stream) {
  return vesicle_stokes<double, false>(src, nullptr, den, qw, src_stride, trg_stride,
                                       n_surfs, trg, head, tail, pot, (hipStream_t)stream);
}
int ves3d_hip_direct_stokes_f32(const float *src, const float *den,
                                const float *qw, size_t src_stride,
                                size_t trg_stride, size_t n_surfs,
                                const float *trg, size_t head, size_t tail,
                                float *pot, ves3d_stream_t stream) {
  return vesicle_stokes<float, false>(src, nullptr, den, qw, src_stride, trg_stride,
                                      n_surfs, trg, head, tail, pot, (hipStream_t)stream);
}
int ves3d_hip_direct_stokes_dl_f64(const double *src, const double *norm,
                                   const double *den, const double *qw,
                                   size_t src_stride, size_t trg_stride,
                                   size_t n_surfs, const double *trg,
                                   size_t head, size_t tail, double *pot,
                                   ves3d_stream_t stream) {
  return vesicle_stokes<double, true>(src, norm, den, qw, src_stride, trg_stride,
                                      n_surfs, trg, head, tail, pot, (hipStream_t)stream);
}
int ves3d_hip_direct_stokes_dl_f32(const float *src, const float *norm,
                                   const float *den, const float *qw,
                                   size_t src_stride, size_t trg_stride,
                                   size_t n_surfs, const float *trg,
                                   size_t head, size_t tail, float *pot,
                                   ves3d_stream_t stream) {
  return vesicle_stokes<float, true>(src, norm, den, qw, src_stride, trg_stride,
                                     n_surfs, trg, head, tail, pot, (hipStream_t)stream);
}

int ves3d_hip_stokes_all2all_f64(const double *src, const double *den,
                                 size_t np, double *pot, ves3d_stream_t stream) {
  return far_eval<false>(src, den, nullptr, np, src, pot, np, 0, 0, 0.0,
                         1.0 / kPi / 8.0, 0.0, (hipStream_t)stream);
}
int ves3d_hip_stokes_all2all_f32c(const double *src, const double *den,
                                  size_t np, double *pot, ves3d_stream_t stream) {
  return far_eval<true>(src, den, nullptr, np, src, pot, np, 0, 0, 0.0,
                        1.0 / kPi / 8.0, 0.0, (hipStream_t)stream);
}

int ves3d_hip_set_far_kernel(int kind) {
  if (kind < 0 || kind > 2) return VES3D_INVALID_PARAMETER;
  g_far_kind.store(kind);
  return VES3D_SUCCESS;
}


int ves3d_hip_direct_eval_f64(const double *src_pos, const double *sl_den,
                              const double *dl_den, size_t n_src,
                              const double *trg_pos, double *trg_vel,
                              size_t n_trg, size_t spg, size_t tpg,
                              ves3d_stream_t stream) {
  return far_eval<false>(src_pos, sl_den, dl_den, n_src, trg_pos, trg_vel, n_trg,
                         spg, tpg, kEvalEps2, 1.0 / (8.0 * kPi),
                         3.0 / (4.0 * kPi), (hipStream_t)stream);
}
int ves3d_hip_direct_eval_f32c(const double *src_pos, const double *sl_den,
                               const double *dl_den, size_t n_src,
                               const double *trg_pos, double *trg_vel,
                               size_t n_trg, size_t spg, size_t tpg,
                               ves3d_stream_t stream) {
  return far_eval<true>(src_pos, sl_den, dl_den, n_src, trg_pos, trg_vel, n_trg,
                        spg, tpg, kEvalEps2, 1.0 / (8.0 * kPi),
                        3.0 / (4.0 * kPi), (hipStream_t)stream);
}

}  // extern "C"

// implicit.hpp -- the globally implicit interfacial-velocity step on the GPU:
// Surface geometry (src/Surface.cc), interfacial forces
// (src/InterfacialForce.cc), the surface Stokes operator (StokesVelocity,
// src/StokesVelocity.cc), the Galerkin matvec / right-hand side / diagonal
// spectral preconditioner (src/InterfacialVelocity.cc:664-1172) and
// BiCGStab (src/BiCGStab.cc:45-138).
#pragma once
#include <vector>

#include "common.hpp"
#include "near.hpp"

namespace v3d {

// Single-layer far field between groups (vesicles): targets of group g see
// the sources of every other group, in FP64 or FP32 with compensated
// accumulation (f32c).  Distinct vesicles' surfaces are disjoint, so the
// pair loop runs without the coincident-point guard of the direct-evaluation
// entry points (stokes_direct.hip).
//
// coarse (optional): one or two coarser sets of sources per group (cpg and
// cpg2 points each, point-major positions and densities) that a wave of 64
// targets uses instead of the fine ones where flags[wave][group] names
// level 1 or 2 (see far_coarse_flags).  The FP32 form stores every level
// relative to the group's first fine source.  With coarse sources the
// scalar-cache kernels run whatever ves3d_hip_set_far_kernel selected.
struct FarCoarse {
  const double *pos;           // level 1: [n_groups * cpg][3]
  const double *den1, *den2;   // densities at the level-1 points (den2: the second of a pair)
  size_t cpg;
  const unsigned char *flags;  // [ceil(n_trg / 64)][n_groups]: the level, 0..2
  // level 2 (cpg2 = 0: none): [n_groups * cpg2][3] points and densities
  const double *pos2 = nullptr;
  const double *l2den1 = nullptr, *l2den2 = nullptr;
  size_t cpg2 = 0;
};
int far_field_between_groups(bool f32c, const double *spos, const double *sld, size_t n_src,
                             const double *tpos, double *tvel, size_t n_trg, size_t spg,
                             size_t tpg, hipStream_t st, const FarCoarse *coarse = nullptr);

// The same with two densities on the same sources and targets, in one pass
// (the pair geometry is computed once and serves both sums).
int far_field_between_groups2(bool f32c, const double *spos, const double *sld1,
                              const double *sld2, size_t n_src, const double *tpos, double *tvel1,
                              double *tvel2, size_t n_trg, size_t spg, size_t tpg, hipStream_t st,
                              const FarCoarse *coarse = nullptr);

// Single- and double-layer far field (the PVFMMEval stand-in's kernels):
// sld [n_src][3] and/or dld [n_src][6] = (density times weight, normal),
// grouped (spg, tpg > 0: own-group pairs skipped) or not; guard: pairs
// closer than sqrt(128 eps) are skipped (targets that may sit on a source).
int far_field_sl_dl(bool f32c, const double *spos, const double *sld, const double *dld,
                    size_t n_src, const double *tpos, double *tvel, size_t n_trg, size_t spg,
                    size_t tpg, bool guard, hipStream_t st);

// ves3d_hip_direct_eval_f64 with caller-provided scratch for the per-slice
// partial sums (part, part_bytes; at least n_trg * 24 bytes, else
// VES3D_INVALID_PARAMETER): allocates nothing.
int direct_eval_f64_fixed(const double *src_pos, const double *sl_den, const double *dl_den,
                          size_t n_src, const double *trg_pos, double *trg_vel, size_t n_trg,
                          size_t spg, size_t tpg, double *part, size_t part_bytes,
                          hipStream_t st);

// flags[w][g] = 2 where every target of wave w (targets 64 w .. 64 w + 63)
// is at least kappa2 R_g from c_g, else 1 where at least kappa R_g, else 0;
// (c_g, R_g) the bounding sphere of group g's fine sources (centre: their
// mean).  flags: ceil(n_trg / 64) * n_groups bytes.  kappa > 1; kappa2 = 0
// (no level 2) or kappa2 > kappa.  counts (device, 3 values): the (wave,
// group) pairs that read level 1, those that read the fine sources (a wave
// skips a group that holds all its targets), and those that read level 2.
int far_coarse_flags(const double *spos, size_t n_src, size_t spg, const double *tpos,
                     size_t n_trg, size_t tpg, double kappa, double kappa2, unsigned char *flags,
                     unsigned long long *counts, hipStream_t st);

// Device buffer owned by the step objects.
struct DBuf {
  double *p = nullptr;
  long n = 0;
  int alloc(long count);
  void release();
  ~DBuf() { release(); }
  operator double *() const { return p; }
};

// Surface<Sca, Vec> at order p: first and second fundamental forms,
// grad / div / linearised mean curvature with the low-pass filter of
// frequency f (src/Surface.cc:211-377, :529-566).
class SurfaceG {
 public:
  int init(int p, int filter_freq, long nv);
  int update(const double *X, hipStream_t st);  // X: [nv][3][ng]
  int grad(const double *f, double *out, hipStream_t st);
  int div(const double *V, double *out, hipStream_t st);
  int lin_mean_curv(const double *x, double *out, hipStream_t st);
  int low_pass(double *field, long nf, hipStream_t st);  // in place, [nf][ng]
  int forward(const double *X, long nf, double *S, hipStream_t st);
  int backward(const double *S, long nf, int du, int dv, double *X, hipStream_t st);
  // The same operators from the input's SH coefficients (a chain of them
  // skips each synthesis that the next operator would transform back):
  // c holds the coefficients of the input field (nv functions for grad,
  // 3 nv for div and lin_mean_curv).  low_pass_c leaves the filtered
  // coefficients in coef and, with want_grid, the filtered field in field;
  // grad_c leaves the low-passed gradient's coefficients in cout (3 nv) and,
  // with want_grid, the gradient in out; lin_mean_curv_c leaves its
  // low-passed result's coefficients in cout (nv).
  int low_pass_c(double *field, long nf, double *coef, bool want_grid, hipStream_t st);
  int grad_c(const double *c, double *out, double *cout, bool want_grid, hipStream_t st);
  int div_c(const double *c, double *out, hipStream_t st);
  // the divergence's low-passed coefficients [nv][nc] (out: grid scratch)
  int div_coef(const double *c, double *out, double *cout, hipStream_t st);
  int lin_mean_curv_c(const double *c, double *out, double *cout, hipStream_t st);

  int p = 0, f = 0;
  long nv = 0, ng = 0, nc = 0;
  DBuf E, F, G, W, cu, cv, nrm, H, K;
  DBuf shc, s1, work, box;
  DBuf tt;  // t1, t2, t3: consecutive [3 nv][ng] slices (grouped syntheses write all three)
  double *t1 = nullptr, *t2 = nullptr, *t3 = nullptr;
  // backward of S (nf functions) for nvar (du, dv) pairs in one grouped
  // synthesis, into t1, t2, t3 in order
  int backward_multi(const double *S, long nf, int nvar, const int (*d)[2], hipStream_t st);
  DBuf c1, c3;  // coefficient buffers of the chains: nv and 3 nv functions
};

// InterfacialForce at order p with the surface upsampled to p_up.
class ForcesG {
 public:
  int init(int p, int p_up, int filter_freq, long nv, double kb);
  int update(const double *X, hipStream_t st);
  int bending(double *Fb, hipStream_t st);                            // explicit
  int linear_bending(const double *x_new, double *Fb, hipStream_t st);
  int tensile(const double *sigma, double *Fs, hipStream_t st);
  // the same from the order-p coefficients [nf][nc] of x_new / sigma
  int linear_bending_c(const double *c, double *Fb, hipStream_t st);
  int tensile_c(const double *c, double *Fs, hipStream_t st);
  // F += rho_v (g . x) n on the p grid (InterfacialForce::gravityForce,
  // src/InterfacialForce.cc:126-160); nothing unless some vesicle has an
  // excess density (VesicleProperties::has_excess_density)
  int gravity_add(const double *x, double *F, hipStream_t st);
  // per-vesicle properties (VesicleProperties, src/VesicleProps.cc:19-55):
  // bending moduli and excess densities [nv] on the host (NULL: unchanged)
  // and the gravity field g[3] (NULL: unchanged)
  int set_props(const double *kb_host, const double *rho_host, const double *g_host);
  // resample [nf][grid(p_from)] -> [nf][grid(p_to)]
  int resample(const double *in, long nf, int p_from, int p_to, double *out, hipStream_t st);
  // the SH coefficients at p_up of a field on the p grid (nf functions):
  // its order-p coefficients, zero-padded (Resample up, then the forward
  // transform at p_up, without the round trip through the upsampled grid)
  int up_coef(const double *in, long nf, double *cu, hipStream_t st);
  // the same from the order-p coefficients [nf][nc]
  int up_coef_c(const double *c, long nf, double *cu, hipStream_t st);

  SurfaceG S, Su;
  int p = 0, pu = 0;
  long nv = 0;
  DBuf bcoef;             // bending coefficient per vesicle, -kappa_b (src/VesicleProps.cc:41-43)
  DBuf rho;               // excess density per vesicle
  double grav[3] = {0.0, 0.0, -1.0};  // Parameters::gravity_field default (src/Parameters.cc:31-33)
  bool has_rho = false;
  DBuf xu, gu, fu, su1, su2, rcoef, rwork;
  DBuf padmap;  // p_up coefficient -> order-p coefficient index, or -1
};

// StokesVelocity<Real>::operator() for surface targets, single layer, no
// viscosity contrast: self (singular operator) + far field (direct sum over
// the other vesicles' upsampled sources).  precision 0: FP64 everywhere;
// 1: FP32 singular operator and FP32-compensated far field; 2: FP64
// singular operator and FP32-compensated far field.
class StokesG {
 public:
  // near: 1 applies the near-singular correction (src/NearSingular.cc) with
  // repulsion length repul_dist (<= 0: none) and near-zone radius r_near
  // (<= 0: the reference's r_ves 2/sqrt(p_up)).
  // far_coarse > 1: the far field reads a vesicle's p-grid sources (its
  // own quadrature) instead of its upsampled ones for every wave of 64
  // targets at least far_coarse times the vesicle's radius from its centre;
  // 0: the upsampled sources for every target (the reference's far field).
  // far_coarse2 > far_coarse (needs far_coarse): beyond far_coarse2 radii
  // the wave reads the vesicle at order p/2 instead: the degree-p surface
  // at the order-p/2 grid points, with the density f times the area element
  // (per unit-sphere measure) truncated to degree p/2 and the order-p/2
  // Gauss-Legendre weights; 0: two levels.
  int init(int p, int p_up, long nv, int precision, int near = 1, double repul_dist = 5e-2,
           double r_near = 0.0, double far_coarse = 0.0, double far_coarse2 = 0.0);
  int setup(const double *X, hipStream_t st);
  // add_repul: the density is f + NearSingular::ForceRepul (the right-hand
  // side of the implicit step, src/InterfacialVelocity.cc:733).  local: the
  // self-interaction (singular operator) alone, without the far field and
  // the near-singular correction.
  // d (optional): a double-layer density [nv][3][ng] (SetDensityDL); f may
  // then be null (no single layer).  The double layer needs enable_dl() and
  // runs over the upsampled sources whatever far_coarse says.
  int apply(const double *f, double *vel, hipStream_t st, bool add_repul = false,
            bool local = false, const double *d = nullptr);
  // The double-layer self operators (built at the next setup) and buffers.
  int enable_dl();
  // Velocity at n targets off the surfaces (trg [n][3] -> vel [n][3];
  // SetTrgCoord, src/StokesVelocity.cc:146-150): the direct sum over every
  // vesicle's upsampled sources (coincident points skipped) with the
  // near-singular correction of targets within r_near of a vesicle, no
  // self term; f (SL) and d (DL) as in apply, either may be null.  FP64.
  int eval_off(const double *f, const double *d, const double *trg, long n, double *vel,
               hipStream_t st);
  // apply for two densities at once (one far-field pass over both): the
  // right-hand side's (f1, with the repulsion when add_repul1) and the
  // solver's first matvec's (f2)
  int apply2(const double *f1, double *vel1, bool add_repul1, const double *f2, double *vel2,
             hipStream_t st);
  // the self interaction alone in SH coefficients: Vc [3 nv][nc] = S_self f
  // (apply with local, without the synthesis of the velocity)
  int self_coef(const double *f, double *Vc, hipStream_t st);

  // HIP events around every far-field launch (on the launch stream); the
  // elapsed times are summed by far_time() after the stream has drained:
  // one-density launches into (ms, calls), two-density ones into (ms2, calls2)
  void timing_begin();
  // near-singular corrections (kind 3) into (ms_near, calls_near)
  int far_time(double *ms, int *calls, double *ms2 = nullptr, int *calls2 = nullptr,
               double *ms_near = nullptr, int *calls_near = nullptr);

  int p = 0, pu = 0, precision = 0;
  long nv = 0;
  bool timing = false;
  std::vector<hipEvent_t> ev0, ev1;
  std::vector<int> evkind;
  int nev = 0;
  DBuf SL;            // FP64 operator (precision 0)
  float *SL32 = nullptr;
  DBuf far_src, trg, aqw, F, V, up, q, farv, shc, work, tmpx, tmpt, tmpp, pole;
  bool use_near = true;
  NearG near;
  DBuf frep, vup, vpole;
  DBuf F2, V2, up2, q2, farv2;  // second density of apply2
  // coarse far-field sources: the p-grid targets with their quadrature
  // weights aqwp (|x_t x x_p| w(lat)), densities qc, qc2, wave flags
  double far_coarse = 0.0;
  DBuf aqwp, qc, qc2;
  unsigned char *cflags = nullptr;
  // device: level-1, fine and level-2 (wave, vesicle) pairs
  unsigned long long *ccounts = nullptr;
  // level 2 (far_coarse2): order ord2 = p/2, its points c2pos [nv][ng2][3],
  // the area element per unit-sphere measure c2gw [nv][ng] on the p grid,
  // densities c2den, c2den2 [nv][ng2][3]; scratch for the product f c2gw
  // [3 nv][ng], its coefficients [3 nv][nc], their order-ord2 synthesis
  // [3 nv][ng2]; the degree mask [nc] and the latitude weights [ord2 + 1]
  double far_coarse2 = 0.0;
  int ord2 = 0;
  DBuf c2pos, c2gw, c2den, c2den2, c2prod, c2coef, c2grid, c2mask, c2w;
  // source-target pairs one far-field launch evaluates (per density), as of
  // the last setup
  double far_pairs = 0;
  int pack_coarse(const double *f, double *qcoarse, hipStream_t st);
  int pack_level2(const double *f, double *out, hipStream_t st);
  // the far-field levels of apply / apply2 (den2 / l2den2: the second density)
  FarCoarse coarse_levels(const double *f1, const double *f2, hipStream_t st, int *err);
  // double layer: operators, upsampled unit normals [nv][3][mu], the
  // density's coefficients G, its packed far-field form qd [nv][M][6],
  // itself on the upsampled grid and at the poles
  bool want_dl = false;
  DBuf DL, nrm_up, G, qd, dup, dpole, selfv;
  float *DL32 = nullptr;
  NearG near_off;              // near-singular correction of targets off the surfaces
  bool near_off_ready = false;
  // The singular operators are built on a stream of their own (bst,
  // non-blocking, default priority), ordered after the setup stream's
  // earlier work (ev_in), so that the far-field pass that needs no self
  // interaction (the right-hand side's) overlaps the build; the first self
  // apply after a setup makes its stream wait for ev_built.
  hipStream_t bst = nullptr;
  hipEvent_t ev_in = nullptr, ev_built = nullptr;
  bool build_pending = false;
  int wait_build(hipStream_t st);
  DBuf dqw, dqwp, dinvs;  // quadrature weights, uploaded at the first setup
  ~StokesG();

 private:
  // density SHTs, self apply (vel = S_self f + D_self d on the p grid) and,
  // with pack, the upsampled far-field densities qc (and qd, dup, dpole)
  int self_and_density(const double *f, double *Fc, double *Vc, double *vel, double *upc,
                       double *qc, bool pack, hipStream_t st, const double *d = nullptr);
  // its three parts: the densities' coefficients (Fc, G), the self velocity
  // (waits for the build), the far-field packing
  int density_coef(const double *f, double *Fc, hipStream_t st, const double *d);
  int self_velocity(bool sl, const double *Fc, double *Vc, double *vel, bool pack,
                    hipStream_t st, bool dl);
  int pack_density(bool sl, const double *Fc, double *upc, double *qc, hipStream_t st, bool dl);
  int near_correct(const double *qc, const double *Vc, double *vel, hipStream_t st,
                   bool dl = false);
  int event_begin(int kind, hipStream_t st);
  int event_end(hipStream_t st);
};

}  // namespace v3d

// implicit.hip -- Surface geometry, interfacial forces and the surface
// Stokes operator of the implicit step (see implicit.hpp).  All fields are
// AxisMajor: vectors [nv][3][ng], scalars [nv][ng]; coefficient arrays are
// function-major [nf][p(p+2)] (the order-major Galerkin layout is produced
// only when packing the solver vector).
#include <algorithm>
#include <cmath>
#include <vector>

#include "implicit.hpp"
#include "runtime.hpp"
#include "sh_tables.hpp"
#include "sht.hpp"
#include "singular.hpp"

extern "C" int ves3d_hip_direct_eval_f64(const double *, const double *, const double *, size_t,
                                         const double *, double *, size_t, size_t, size_t, void *);
extern "C" int ves3d_hip_direct_eval_f32c(const double *, const double *, const double *, size_t,
                                          const double *, double *, size_t, size_t, size_t, void *);

namespace v3d {

int DBuf::alloc(long count) {
  release();
  if (count <= 0) count = 1;
  if (hipMalloc(&p, sizeof(double) * count) != hipSuccess) {
    p = nullptr;
    return VES3D_MEMORY_ERROR;
  }
  n = count;
  return VES3D_SUCCESS;
}
void DBuf::release() {
  if (p) (void)hipFree(p);
  p = nullptr;
  n = 0;
}

namespace {
inline unsigned gblocks(long n) {
  long b = (n + 255) / 256;
  if (b > 65536) b = 65536;
  return (unsigned)(b < 1 ? 1 : b);
}
#define GRID_LOOP(i, n) \
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < (n); i += (long)gridDim.x * blockDim.x)

// First fundamental forms, area element, normal and the grad/div
// coefficients cu, cv (src/Surface.cc:211-262).
__global__ void k_first_forms(const double *__restrict__ du, const double *__restrict__ dv,
                              long nv, long ng, double *E, double *F, double *G, double *W,
                              double *cu, double *cv, double *nrm) {
  GRID_LOOP(i, nv * ng) {
    const long v = i / ng, g = i % ng;
    const long b = v * 3 * ng + g;
    const double ux = du[b], uy = du[b + ng], uz = du[b + 2 * ng];
    const double vx = dv[b], vy = dv[b + ng], vz = dv[b + 2 * ng];
    const double e = ux * ux + uy * uy + uz * uz;
    const double f = ux * vx + uy * vy + uz * vz;
    const double gg = vx * vx + vy * vy + vz * vz;
    const double w2 = e * gg - f * f;
    const double es = e / w2, fs = f / w2, gs = gg / w2;
    const double w = sqrt(w2);
    E[i] = es;
    F[i] = fs;
    G[i] = gs;
    W[i] = w;
    cu[b] = gs * ux - fs * vx;
    cu[b + ng] = gs * uy - fs * vy;
    cu[b + 2 * ng] = gs * uz - fs * vz;
    cv[b] = es * vx - fs * ux;
    cv[b + ng] = es * vy - fs * uy;
    cv[b + 2 * ng] = es * vz - fs * uz;
    nrm[b] = (uy * vz - uz * vy) / w;
    nrm[b + ng] = (uz * vx - ux * vz) / w;
    nrm[b + 2 * ng] = (ux * vy - uy * vx) / w;
  }
}

// Mean and Gaussian curvature before filtering (src/Surface.cc:264-321).
__global__ void k_second_forms(const double *__restrict__ duv, const double *__restrict__ d2u,
                               const double *__restrict__ d2v, const double *__restrict__ nrm,
                               const double *E, const double *F, const double *G,
                               const double *W, long nv, long ng, double *H, double *K) {
  GRID_LOOP(i, nv * ng) {
    const long v = i / ng, g = i % ng;
    const long b = v * 3 * ng + g;
    const double nx = nrm[b], ny = nrm[b + ng], nz = nrm[b + 2 * ng];
    const double M = duv[b] * nx + duv[b + ng] * ny + duv[b + 2 * ng] * nz;
    const double L = d2u[b] * nx + d2u[b + ng] * ny + d2u[b + 2 * ng] * nz;
    const double N = d2v[b] * nx + d2v[b + ng] * ny + d2v[b + 2 * ng] * nz;
    H[i] = -F[i] * M + 0.5 * G[i] * L + 0.5 * E[i] * N;
    K[i] = (L * N - M * M) / W[i] / W[i];
  }
}

// Linearised mean curvature before filtering (src/Surface.cc:529-566).
__global__ void k_lin_mean_curv(const double *__restrict__ d2u, const double *__restrict__ duv,
                                const double *__restrict__ d2v, const double *__restrict__ nrm,
                                const double *E, const double *F, const double *G, long nv,
                                long ng, double *out) {
  GRID_LOOP(i, nv * ng) {
    const long v = i / ng, g = i % ng;
    const long b = v * 3 * ng + g;
    const double nx = nrm[b], ny = nrm[b + ng], nz = nrm[b + 2 * ng];
    const double a = d2u[b] * nx + d2u[b + ng] * ny + d2u[b + 2 * ng] * nz;
    const double m = duv[b] * nx + duv[b + ng] * ny + duv[b + 2 * ng] * nz;
    const double c = d2v[b] * nx + d2v[b + ng] * ny + d2v[b + 2 * ng] * nz;
    out[i] = 0.5 * (G[i] * a - 2.0 * F[i] * m + E[i] * c);
  }
}

// grad f = f_u cu + f_v cv  (src/Surface.cc:323-346)
__global__ void k_grad_comb(const double *__restrict__ fu, const double *__restrict__ fv,
                            const double *__restrict__ cu, const double *__restrict__ cv,
                            long nv, long ng, double *out) {
  GRID_LOOP(i, nv * ng) {
    const long v = i / ng, g = i % ng;
    const long b = v * 3 * ng + g;
    const double a = fu[i], c = fv[i];
#pragma unroll
    for (int d = 0; d < 3; ++d) out[b + d * ng] = a * cu[b + d * ng] + c * cv[b + d * ng];
  }
}

// div V = V_u . cu + V_v . cv  (src/Surface.cc:348-377)
__global__ void k_div_comb(const double *__restrict__ Vu, const double *__restrict__ Vv,
                           const double *__restrict__ cu, const double *__restrict__ cv, long nv,
                           long ng, double *out) {
  GRID_LOOP(i, nv * ng) {
    const long v = i / ng, g = i % ng;
    const long b = v * 3 * ng + g;
    double s = 0;
#pragma unroll
    for (int d = 0; d < 3; ++d) s += Vv[b + d * ng] * cv[b + d * ng];
    double t = 0;
#pragma unroll
    for (int d = 0; d < 3; ++d) t += Vu[b + d * ng] * cu[b + d * ng];
    out[i] = s + t;
  }
}

__global__ void k_scale_rows(double *S, long nf, int nc, const double *__restrict__ scale) {
  GRID_LOOP(i, nf * nc) S[i] *= scale[i % nc];
}

// s2 = a * (H^2 - K)
__global__ void k_bend_s2(const double *__restrict__ a, const double *__restrict__ H,
                          const double *__restrict__ K, long n, double *s2) {
  GRID_LOOP(i, n) s2[i] = a[i] * (H[i] * H[i] - K[i]);
}

// F = c_v (s1 + 2 s2) n with c_v vesicle v's bending coefficient, -kappa_b
// (src/InterfacialForce.cc:37-41, src/VesicleProps.cc:41-43)
__global__ void k_bend_force(const double *__restrict__ s1, const double *__restrict__ s2,
                             const double *__restrict__ nrm, const double *__restrict__ coeff,
                             long nv, long ng, double *Fo) {
  GRID_LOOP(i, nv * ng) {
    const long v = i / ng, g = i % ng;
    const long b = v * 3 * ng + g;
    const double s = coeff[v] * (s1[i] + 2.0 * s2[i]);
#pragma unroll
    for (int d = 0; d < 3; ++d) Fo[b + d * ng] = s * nrm[b + d * ng];
  }
}

// F += rho_v (g . x) n   (src/InterfacialForce.cc:140-159)
__global__ void k_gravity_add(const double *__restrict__ x, const double *__restrict__ nrm,
                              const double *__restrict__ rho, double gx, double gy, double gz,
                              long nv, long ng, double *F) {
  GRID_LOOP(i, nv * ng) {
    const long v = i / ng, g = i % ng;
    const long b = v * 3 * ng + g;
    const double s = rho[v] * (gx * x[b] + gy * x[b + ng] + gz * x[b + 2 * ng]);
#pragma unroll
    for (int d = 0; d < 3; ++d) F[b + d * ng] += s * nrm[b + d * ng];
  }
}

// F = 2 sigma H n + grad sigma   (src/InterfacialForce.cc:109-112)
__global__ void k_tens_force(const double *__restrict__ sig, const double *__restrict__ H,
                             const double *__restrict__ nrm, const double *__restrict__ gs,
                             long nv, long ng, double *Fo) {
  GRID_LOOP(i, nv * ng) {
    const long v = i / ng, g = i % ng;
    const long b = v * 3 * ng + g;
    const double s = 2.0 * sig[i] * H[i];
#pragma unroll
    for (int d = 0; d < 3; ++d) Fo[b + d * ng] = s * nrm[b + d * ng] + gs[b + d * ng];
  }
}

// Far-field source setup (src/StokesVelocity.cc:196-277, :307-337):
// far[v][0..1] = poles, far[v][2+j] = upsampled points; aqw = |x_t x x_p| w(lat).
// nrm (optional): the unit normals (x_t x x_p) / |x_t x x_p|, AxisMajor
// [nv][3][mu] (scoord_norm, src/StokesVelocity.cc:244-273).
__global__ void k_far_setup(const double *__restrict__ X, const double *__restrict__ Xt,
                            const double *__restrict__ Xp, const double *__restrict__ pole,
                            const double *__restrict__ qw, long nv, int pu, double *far,
                            double *aqw, double *nrm) {
  const long mu = 2L * pu * (pu + 1);
  GRID_LOOP(i, nv * (mu + 2)) {
    const long v = i / (mu + 2), j = i % (mu + 2);
    double *o = far + i * 3;
    if (j < 2) {
#pragma unroll
      for (int d = 0; d < 3; ++d) o[d] = pole[(v * 3 + d) * 2 + j];
    } else {
      const long s = j - 2;
      const long b = v * 3 * mu + s;
#pragma unroll
      for (int d = 0; d < 3; ++d) o[d] = X[b + d * mu];
      const double xt = Xt[b], yt = Xt[b + mu], zt = Xt[b + 2 * mu];
      const double xp = Xp[b], yp = Xp[b + mu], zp = Xp[b + 2 * mu];
      const double nx = yt * zp - zt * yp, ny = zt * xp - xt * zp, nz = xt * yp - yt * xp;
      const double area = sqrt(nx * nx + ny * ny + nz * nz);
      aqw[v * mu + s] = area * qw[s / (2 * pu)];
      if (nrm) {
        const double ia = 1.0 / area;
        nrm[b] = nx * ia;
        nrm[b + mu] = ny * ia;
        nrm[b + 2 * mu] = nz * ia;
      }
    }
  }
}

// qd[v][0..1] = 0, qd[v][2+s] = (d_up[v][:][s] aqw[v][s], n_up[v][:][s])
// (qforce_double, src/StokesVelocity.cc:338-378)
__global__ void k_pack_dl(const double *__restrict__ du, const double *__restrict__ aqw,
                          const double *__restrict__ nrm, long nv, long mu, double *qd) {
  GRID_LOOP(i, nv * (mu + 2)) {
    const long v = i / (mu + 2), j = i % (mu + 2);
    double *o = qd + i * 6;
    if (j < 2) {
#pragma unroll
      for (int d = 0; d < 6; ++d) o[d] = 0.0;
    } else {
      const long s = j - 2;
      const double w = aqw[v * mu + s];
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        o[d] = du[(v * 3 + d) * mu + s] * w;
        o[3 + d] = nrm[(v * 3 + d) * mu + s];
      }
    }
  }
}

// p-grid quadrature weight of the coarse far-field sources:
// aqwp[v][g] = |x_t x x_p| w(lat) (as k_far_setup on the upsampled grid)
__global__ void k_coarse_weights(const double *__restrict__ Xt, const double *__restrict__ Xp,
                                 const double *__restrict__ qw, long nv, int p,
                                 double *__restrict__ aqwp) {
  const long ng = 2L * p * (p + 1);
  GRID_LOOP(i, nv * ng) {
    const long v = i / ng, s = i % ng;
    const long b = v * 3 * ng + s;
    const double xt = Xt[b], yt = Xt[b + ng], zt = Xt[b + 2 * ng];
    const double xp = Xp[b], yp = Xp[b + ng], zp = Xp[b + 2 * ng];
    const double nx = yt * zp - zt * yp, ny = zt * xp - xt * zp, nz = xt * yp - yt * xp;
    aqwp[i] = sqrt(nx * nx + ny * ny + nz * nz) * qw[s / (2 * p)];
  }
}

// out[f][k] = in[f][map[k]] (map[k] >= 0) or 0: zero-padding of SH
// coefficients to a higher order
__global__ void k_pad_coef(const double *__restrict__ in, long nf, long nci,
                           const double *__restrict__ map, long nco, double *__restrict__ out) {
  GRID_LOOP(i, nf * nco) {
    const long f = i / nco;
    const double m = map[i % nco];
    out[i] = m >= 0.0 ? in[f * nci + (long)m] : 0.0;
  }
}
// y = float(x), round to nearest
__global__ void k_round_f32(const double *__restrict__ x, long n, float *__restrict__ y) {
  GRID_LOOP(i, n) y[i] = (float)x[i];
}
// out[v][d][g] = f[v][d][g] * w[v][g]
__global__ void k_mul_pointwise(const double *__restrict__ f, const double *__restrict__ w, long nv,
                                long ng, double *__restrict__ out) {
  GRID_LOOP(i, 3 * nv * ng) out[i] = f[i] * w[(i / (3 * ng)) * ng + i % ng];
}
// c[f][k] *= mask[k] (zeroes the coefficients above a degree)
__global__ void k_mask_coef(double *__restrict__ c, const double *__restrict__ mask, long n,
                            long nc) {
  GRID_LOOP(i, n) c[i] *= mask[i % nc];
}
// out[v ng2 + i][d] = grid[v][d][i] * w[latitude of i] (order-q2 grid,
// 2 q2 longitudes per latitude)
__global__ void k_pack_level2(const double *__restrict__ grid, const double *__restrict__ w,
                              long nv, long ng2, int q2, double *__restrict__ out) {
  GRID_LOOP(i, nv * ng2) {
    const long v = i / ng2, k = i % ng2;
    const double wk = w[k / (2 * q2)];
#pragma unroll
    for (int d = 0; d < 3; ++d) out[i * 3 + d] = grid[(v * 3 + d) * ng2 + k] * wk;
  }
}
// qc[v ng + g][d] = f[v][d][g] * aqwp[v][g] (point-major coarse densities)
__global__ void k_pack_coarse(const double *__restrict__ f, const double *__restrict__ aqwp,
                              long nv, long ng, double *__restrict__ qc) {
  GRID_LOOP(i, nv * ng) {
    const long v = i / ng, g = i % ng;
    const double w = aqwp[i];
#pragma unroll
    for (int d = 0; d < 3; ++d) qc[i * 3 + d] = f[(v * 3 + d) * ng + g] * w;
  }
}

// AxisMajor [nv][3][ng] -> point-major [nv*ng][3]
__global__ void k_to_points(const double *__restrict__ X, long nv, long ng, double *P) {
  GRID_LOOP(i, nv * ng) {
    const long v = i / ng, g = i % ng;
#pragma unroll
    for (int d = 0; d < 3; ++d) P[i * 3 + d] = X[(v * 3 + d) * ng + g];
  }
}

// q[v][0..1] = 0, q[v][2+s] = f_up[v][:][s] * aqw[v][s]
__global__ void k_pack_density(const double *__restrict__ fu, const double *__restrict__ aqw,
                               long nv, long mu, double *q) {
  GRID_LOOP(i, nv * (mu + 2)) {
    const long v = i / (mu + 2), j = i % (mu + 2);
    double *o = q + i * 3;
    if (j < 2) {
      o[0] = o[1] = o[2] = 0.0;
    } else {
      const long s = j - 2;
      const double w = aqw[v * mu + s];
#pragma unroll
      for (int d = 0; d < 3; ++d) o[d] = fu[(v * 3 + d) * mu + s] * w;
    }
  }
}

// vel[v][d][g] += far[v*ng+g][d]
__global__ void k_add_points(const double *__restrict__ P, long nv, long ng, double *vel) {
  GRID_LOOP(i, nv * ng) {
    const long v = i / ng, g = i % ng;
#pragma unroll
    for (int d = 0; d < 3; ++d) vel[(v * 3 + d) * ng + g] += P[i * 3 + d];
  }
}

std::vector<double> box_filter(int p, int f) {  // src/SHTrans.cc:23-31
  std::vector<double> b(sh_ncoef(p));
  for (int j = 0; j < 2 * p; ++j) {
    const int m = (j + 1) / 2, off = sh_blockoff(p, j), len = sh_blocklen(p, j);
    for (int k = 0; k < len; ++k) b[off + k] = (m + k) > f ? 0.0 : 1.0;
  }
  return b;
}

int upload(DBuf &d, const std::vector<double> &h) {
  V3D_CHECK(d.alloc((long)h.size()));
  return hip_status(hipMemcpy(d.p, h.data(), sizeof(double) * h.size(), hipMemcpyHostToDevice));
}
}  // namespace

// ------------------------------------------------------------- SurfaceG
int SurfaceG::init(int p_, int f_, long nv_) {
  p = p_;
  f = f_;
  nv = nv_;
  ng = sh_ngrid(p);
  nc = sh_ncoef(p);
  for (DBuf *b : {&E, &F, &G, &W, &H, &K, &s1}) V3D_CHECK(b->alloc(nv * ng));
  for (DBuf *b : {&cu, &cv, &nrm}) V3D_CHECK(b->alloc(3 * nv * ng));
  V3D_CHECK(tt.alloc(9 * nv * ng));
  t1 = tt.p;
  t2 = tt.p + 3 * nv * ng;
  t3 = tt.p + 6 * nv * ng;
  V3D_CHECK(shc.alloc(3 * nv * nc));
  V3D_CHECK(c1.alloc(nv * nc));
  V3D_CHECK(c3.alloc(3 * nv * nc));
  V3D_CHECK(work.alloc(
      std::max(shc2grid_multi_work(3 * nv, p, p, 3), grid2shc_work(3 * nv, p, p))));
  return upload(box, box_filter(p, f));
}

int SurfaceG::forward(const double *X, long nf, double *S, hipStream_t st) {
  return sht_grid2shc<double>(X, nf, p, p, S, work, st);
}

// SHTrans::backward / backward_du / dv / d2u / d2v / duv: du, dv = number of
// latitude / longitude derivatives.  SHTrans's first latitude derivative is
// the negative of SphericalHarmonics' LegPolyDeriv (d1legTrans files).
int SurfaceG::backward(const double *S, long nf, int du, int dv, double *X, hipStream_t st) {
  return sht_synth<double>(S, nf, p, p, du, dv, X, work, st, du == 1 ? -1.0 : 1.0);
}

int SurfaceG::backward_multi(const double *S, long nf, int nvar, const int (*d)[2],
                             hipStream_t st) {
  int du[3], dv[3], neg[3];
  for (int v = 0; v < nvar; ++v) {
    du[v] = d[v][0];
    dv[v] = d[v][1];
    neg[v] = du[v] == 1;  // SHTrans's first latitude derivative (see backward)
  }
  return sht_synth_multi<double>(S, nf, p, p, nvar, du, dv, neg, t1, 3 * nv * ng, work, st);
}

int SurfaceG::low_pass_c(double *field, long nf, double *coef, bool want_grid, hipStream_t st) {
  V3D_CHECK(forward(field, nf, coef, st));
  k_scale_rows<<<gblocks(nf * nc), 256, 0, st>>>(coef, nf, (int)nc, box);
  V3D_CHECK(launch_status());
  return want_grid ? backward(coef, nf, 0, 0, field, st) : VES3D_SUCCESS;
}

int SurfaceG::low_pass(double *field, long nf, hipStream_t st) {
  return low_pass_c(field, nf, shc, true, st);
}

int SurfaceG::update(const double *X, hipStream_t st) {
  V3D_CHECK(forward(X, 3 * nv, shc, st));
  static const int d1[2][2] = {{1, 0}, {0, 1}}, d2[3][2] = {{1, 1}, {2, 0}, {0, 2}};
  V3D_CHECK(backward_multi(shc, 3 * nv, 2, d1, st));
  k_first_forms<<<gblocks(nv * ng), 256, 0, st>>>(t1, t2, nv, ng, E, F, G, W, cu, cv, nrm);
  V3D_CHECK(launch_status());
  V3D_CHECK(backward_multi(shc, 3 * nv, 3, d2, st));
  k_second_forms<<<gblocks(nv * ng), 256, 0, st>>>(t1, t2, t3, nrm, E, F, G, W, nv, ng, H, K);
  V3D_CHECK(launch_status());
  V3D_CHECK(low_pass(K, nv, st));
  return low_pass(H, nv, st);
}

static const int kD1[2][2] = {{1, 0}, {0, 1}};

int SurfaceG::grad_c(const double *c, double *out, double *cout, bool want_grid,
                     hipStream_t st) {
  V3D_CHECK(backward_multi(c, nv, 2, kD1, st));
  k_grad_comb<<<gblocks(nv * ng), 256, 0, st>>>(t1, t2, cu, cv, nv, ng, out);
  V3D_CHECK(launch_status());
  return low_pass_c(out, 3 * nv, cout, want_grid, st);
}

int SurfaceG::grad(const double *fin, double *out, hipStream_t st) {
  V3D_CHECK(forward(fin, nv, shc, st));
  return grad_c(shc, out, c3, true, st);
}

int SurfaceG::div_c(const double *c, double *out, hipStream_t st) {
  V3D_CHECK(backward_multi(c, 3 * nv, 2, kD1, st));
  k_div_comb<<<gblocks(nv * ng), 256, 0, st>>>(t1, t2, cu, cv, nv, ng, out);
  V3D_CHECK(launch_status());
  return low_pass(out, nv, st);
}

int SurfaceG::div_coef(const double *c, double *out, double *cout, hipStream_t st) {
  V3D_CHECK(backward_multi(c, 3 * nv, 2, kD1, st));
  k_div_comb<<<gblocks(nv * ng), 256, 0, st>>>(t1, t2, cu, cv, nv, ng, out);
  V3D_CHECK(launch_status());
  return low_pass_c(out, nv, cout, false, st);
}

int SurfaceG::div(const double *V, double *out, hipStream_t st) {
  V3D_CHECK(forward(V, 3 * nv, shc, st));
  return div_c(shc, out, st);
}

int SurfaceG::lin_mean_curv_c(const double *c, double *out, double *cout, hipStream_t st) {
  static const int d[3][2] = {{2, 0}, {1, 1}, {0, 2}};
  V3D_CHECK(backward_multi(c, 3 * nv, 3, d, st));
  k_lin_mean_curv<<<gblocks(nv * ng), 256, 0, st>>>(t1, t2, t3, nrm, E, F, G, nv, ng, out);
  V3D_CHECK(launch_status());
  return low_pass_c(out, nv, cout, true, st);
}

int SurfaceG::lin_mean_curv(const double *x, double *out, hipStream_t st) {
  V3D_CHECK(forward(x, 3 * nv, shc, st));
  return lin_mean_curv_c(shc, out, c1, st);
}

// ------------------------------------------------------------- ForcesG
int ForcesG::init(int p_, int pu_, int f, long nv_, double kb_) {
  p = p_;
  pu = pu_;
  nv = nv_;
  V3D_CHECK(bcoef.alloc(nv));
  V3D_CHECK(rho.alloc(nv));
  {
    std::vector<double> c(nv, -kb_), z(nv, 0.0);
    V3D_HIP(hipMemcpy(bcoef.p, c.data(), sizeof(double) * nv, hipMemcpyHostToDevice));
    V3D_HIP(hipMemcpy(rho.p, z.data(), sizeof(double) * nv, hipMemcpyHostToDevice));
  }
  V3D_CHECK(S.init(p, f, nv));
  V3D_CHECK(Su.init(pu, f * pu / p, nv));  // filter kept relative (src/Surface.cc:589-590)
  const long mu = sh_ngrid(pu);
  for (DBuf *b : {&xu, &gu, &fu}) V3D_CHECK(b->alloc(3 * nv * mu));
  for (DBuf *b : {&su1, &su2}) V3D_CHECK(b->alloc(nv * mu));
  const int pm = std::min(p, pu);
  V3D_CHECK(rcoef.alloc(3 * nv * sh_ncoef(pm)));
  long w = std::max({grid2shc_work(3 * nv, p, pm), grid2shc_work(3 * nv, pu, pm),
                     shc2grid_work(3 * nv, pm, p), shc2grid_work(3 * nv, pm, pu),
                     grid2shc_work(3 * nv, p, p)});
  if (pu >= p) {
    // block j (order m = (j + 1) / 2) of the p_up layout starts with the
    // degrees m .. p that block j of the order-p layout holds
    std::vector<double> map(sh_ncoef(pu), -1.0);
    for (int j = 0; j < 2 * p; ++j)
      for (int k = 0; k < sh_blocklen(p, j); ++k)
        map[sh_blockoff(pu, j) + k] = sh_blockoff(p, j) + k;
    V3D_CHECK(upload(padmap, map));
  }
  return rwork.alloc(w);
}

int ForcesG::up_coef(const double *in, long nf, double *cu_out, hipStream_t st) {
  if (!padmap.p) return VES3D_INVALID_PARAMETER;
  V3D_CHECK(sht_grid2shc<double>(in, nf, p, p, rcoef, rwork, st));
  return up_coef_c(rcoef, nf, cu_out, st);
}

int ForcesG::up_coef_c(const double *c, long nf, double *cu_out, hipStream_t st) {
  if (!padmap.p) return VES3D_INVALID_PARAMETER;
  const long ncu = sh_ncoef(pu);
  k_pad_coef<<<gblocks(nf * ncu), 256, 0, st>>>(c, nf, sh_ncoef(p), padmap, ncu, cu_out);
  return launch_status();
}

// Resample (src/HelperFuns.cc:436-493): truncate / zero-pad the SH expansion.
int ForcesG::resample(const double *in, long nf, int pf, int pt, double *out, hipStream_t st) {
  if (pf == pt) return hip_status(hipMemcpyAsync(out, in, sizeof(double) * nf * sh_ngrid(pf),
                                                 hipMemcpyDeviceToDevice, st));
  const int pm = std::min(pf, pt);
  V3D_CHECK(sht_grid2shc<double>(in, nf, pf, pm, rcoef, rwork, st));
  return sht_shc2grid<double>(rcoef, nf, pm, pt, out, nullptr, nullptr, rwork, st);
}

int ForcesG::update(const double *X, hipStream_t st) {
  V3D_CHECK(S.update(X, st));
  V3D_CHECK(resample(X, 3 * nv, p, pu, xu, st));
  return Su.update(xu, st);
}

int ForcesG::bending(double *Fb, hipStream_t st) {  // src/InterfacialForce.cc:20-50
  const long mu = Su.ng;
  k_bend_s2<<<gblocks(nv * mu), 256, 0, st>>>(Su.H, Su.H, Su.K, nv * mu, su2);
  V3D_CHECK(launch_status());
  // div(grad H): the low-passed gradient's coefficients feed the divergence
  V3D_CHECK(Su.forward(Su.H, nv, Su.c1, st));
  V3D_CHECK(Su.grad_c(Su.c1, gu, Su.c3, false, st));
  V3D_CHECK(Su.div_c(Su.c3, su1, st));
  k_bend_force<<<gblocks(nv * mu), 256, 0, st>>>(su1, su2, Su.nrm, bcoef, nv, mu, fu);
  V3D_CHECK(launch_status());
  return resample(fu, 3 * nv, pu, p, Fb, st);
}

int ForcesG::linear_bending(const double *x_new, double *Fb, hipStream_t st) {  // :52-92
  V3D_CHECK(sht_grid2shc<double>(x_new, 3 * nv, p, p, rcoef, rwork, st));
  return linear_bending_c(rcoef, Fb, st);
}

int ForcesG::linear_bending_c(const double *c, double *Fb, hipStream_t st) {
  const long mu = Su.ng;
  // x_new upsampled, its linearised mean curvature, the gradient of that
  // and its divergence, each operator fed the previous one's low-passed
  // coefficients (the reference synthesises each on the upsampled grid and
  // transforms it back: the same field, up to rounding)
  V3D_CHECK(up_coef_c(c, 3 * nv, Su.c3, st));
  V3D_CHECK(Su.lin_mean_curv_c(Su.c3, su1, Su.c1, st));
  k_bend_s2<<<gblocks(nv * mu), 256, 0, st>>>(su1, Su.H, Su.K, nv * mu, su2);
  V3D_CHECK(launch_status());
  V3D_CHECK(Su.grad_c(Su.c1, gu, Su.c3, false, st));
  V3D_CHECK(Su.div_c(Su.c3, su1, st));
  k_bend_force<<<gblocks(nv * mu), 256, 0, st>>>(su1, su2, Su.nrm, bcoef, nv, mu, fu);
  V3D_CHECK(launch_status());
  return resample(fu, 3 * nv, pu, p, Fb, st);
}

int ForcesG::set_props(const double *kb_host, const double *rho_host, const double *g_host) {
  if (kb_host) {
    std::vector<double> c(nv);
    for (long v = 0; v < nv; ++v) c[v] = -kb_host[v];  // bending_coeff = -bending_modulus
    V3D_HIP(hipMemcpy(bcoef.p, c.data(), sizeof(double) * nv, hipMemcpyHostToDevice));
  }
  if (rho_host) {
    V3D_HIP(hipMemcpy(rho.p, rho_host, sizeof(double) * nv, hipMemcpyHostToDevice));
    has_rho = false;  // VesicleProperties::update (src/VesicleProps.cc:50-52)
    for (long v = 0; v < nv; ++v) has_rho |= std::fabs(rho_host[v]) > 1e-12;
  }
  if (g_host)
    for (int d = 0; d < 3; ++d) grav[d] = g_host[d];
  return VES3D_SUCCESS;
}

int ForcesG::gravity_add(const double *x, double *F, hipStream_t st) {
  if (!has_rho) return VES3D_SUCCESS;
  const long ng = S.ng;
  k_gravity_add<<<gblocks(nv * ng), 256, 0, st>>>(x, S.nrm, rho, grav[0], grav[1], grav[2], nv,
                                                  ng, F);
  return launch_status();
}

int ForcesG::tensile(const double *sigma, double *Fs, hipStream_t st) {  // :94-124
  V3D_CHECK(sht_grid2shc<double>(sigma, nv, p, p, rcoef, rwork, st));
  return tensile_c(rcoef, Fs, st);
}

int ForcesG::tensile_c(const double *c, double *Fs, hipStream_t st) {
  const long mu = Su.ng;
  // sigma upsampled (from its zero-padded coefficients) and its gradient
  V3D_CHECK(up_coef_c(c, nv, Su.c1, st));
  V3D_CHECK(Su.backward(Su.c1, nv, 0, 0, su1, st));
  V3D_CHECK(Su.grad_c(Su.c1, gu, Su.c3, true, st));
  k_tens_force<<<gblocks(nv * mu), 256, 0, st>>>(su1, Su.H, Su.nrm, gu, nv, mu, fu);
  V3D_CHECK(launch_status());
  return resample(fu, 3 * nv, pu, p, Fs, st);
}

// ------------------------------------------------------------- StokesG
StokesG::~StokesG() {
  if (bst) {  // a build in flight writes SL
    (void)hipStreamSynchronize(bst);
    (void)hipStreamDestroy(bst);
  }
  if (ev_in) (void)hipEventDestroy(ev_in);
  if (ev_built) (void)hipEventDestroy(ev_built);
  if (SL32) (void)hipFree(SL32);
  if (DL32) (void)hipFree(DL32);
  if (cflags) (void)hipFree(cflags);
  if (ccounts) (void)hipFree(ccounts);
  for (auto e : ev0) (void)hipEventDestroy(e);
  for (auto e : ev1) (void)hipEventDestroy(e);
}

void StokesG::timing_begin() {
  timing = true;
  nev = 0;
}

int StokesG::far_time(double *ms, int *calls, double *ms2, int *calls2, double *ms_near,
                      int *calls_near) {
  double tot[4] = {0, 0, 0, 0};
  int cnt[4] = {0, 0, 0, 0};
  for (int i = 0; i < nev; ++i) {
    V3D_HIP(hipEventSynchronize(ev1[i]));
    float t = 0;
    V3D_HIP(hipEventElapsedTime(&t, ev0[i], ev1[i]));
    tot[evkind[i]] += t;
    ++cnt[evkind[i]];
  }
  if (ms) *ms = tot[1];
  if (calls) *calls = cnt[1];
  if (ms2) *ms2 = tot[2];
  if (calls2) *calls2 = cnt[2];
  if (ms_near) *ms_near = tot[3];
  if (calls_near) *calls_near = cnt[3];
  timing = false;
  nev = 0;
  return VES3D_SUCCESS;
}

int StokesG::init(int p_, int pu_, long nv_, int precision_, int near_, double repul_dist,
                  double r_near, double far_coarse_, double far_coarse2_) {
  p = p_;
  pu = pu_;
  nv = nv_;
  precision = precision_;
  use_near = near_ != 0;
  if (far_coarse_ != 0.0 && !(far_coarse_ > 1.0)) return VES3D_INVALID_PARAMETER;
  if (far_coarse2_ != 0.0 && !(far_coarse_ > 1.0 && far_coarse2_ > far_coarse_ && p_ >= 4))
    return VES3D_INVALID_PARAMETER;
  far_coarse = far_coarse_;
  far_coarse2 = far_coarse2_;
  const long nc = sh_ncoef(p), ng = sh_ngrid(p), mu = sh_ngrid(pu);
  const long nmat = nv * 9 * nc * nc;
  if (!bst) {
    V3D_HIP(hipStreamCreateWithFlags(&bst, hipStreamNonBlocking));
    V3D_HIP(hipEventCreateWithFlags(&ev_in, hipEventDisableTiming));
    V3D_HIP(hipEventCreateWithFlags(&ev_built, hipEventDisableTiming));
  }
  if (precision != 1) V3D_CHECK(SL.alloc(nmat));
  // precision 1: the FP32 build; precision 2: an FP32 copy of the FP64
  // operator for the local (inner-solve) applications
  if (precision != 0)
    if (hipMalloc(&SL32, sizeof(float) * nmat) != hipSuccess) return VES3D_MEMORY_ERROR;
  V3D_CHECK(far_src.alloc(nv * (mu + 2) * 3));
  V3D_CHECK(q.alloc(nv * (mu + 2) * 3));
  V3D_CHECK(trg.alloc(nv * ng * 3));
  V3D_CHECK(farv.alloc(nv * ng * 3));
  V3D_CHECK(aqw.alloc(nv * mu));
  for (DBuf *b : {&F, &V, &shc}) V3D_CHECK(b->alloc(3 * nv * nc));
  for (DBuf *b : {&up, &tmpx, &tmpt, &tmpp}) V3D_CHECK(b->alloc(3 * nv * mu));
  V3D_CHECK(pole.alloc(3 * nv * 2));
  if (far_coarse > 0) {
    V3D_CHECK(aqwp.alloc(nv * ng));
    V3D_CHECK(qc.alloc(nv * ng * 3));
    const long nw = (nv * ng + 63) / 64;
    if (hipMalloc(&cflags, (size_t)nw * nv) != hipSuccess) return VES3D_MEMORY_ERROR;
    if (hipMalloc(&ccounts, 3 * sizeof(unsigned long long)) != hipSuccess)
      return VES3D_MEMORY_ERROR;
  }
  if (far_coarse2 > 0) {
    ord2 = p / 2;
    const long ng2 = sh_ngrid(ord2);
    V3D_CHECK(c2pos.alloc(nv * ng2 * 3));
    V3D_CHECK(c2gw.alloc(nv * ng));
    V3D_CHECK(c2den.alloc(nv * ng2 * 3));
    V3D_CHECK(c2prod.alloc(3 * nv * ng));
    V3D_CHECK(c2coef.alloc(3 * nv * nc));
    V3D_CHECK(c2grid.alloc(3 * nv * ng2));
    std::vector<double> mask(nc, 0.0), x, w;
    for (int j = 0; j < 2 * p; ++j) {
      const int m = (j + 1) / 2;
      for (int k = 0; k < sh_blocklen(p, j); ++k) mask[sh_blockoff(p, j) + k] = m + k <= ord2;
    }
    V3D_CHECK(upload(c2mask, mask));
    gauss_legendre(ord2 + 1, x, w);
    for (double &v : w) v *= kPi / ord2;
    V3D_CHECK(upload(c2w, w));
  }
  if (use_near) {
    V3D_CHECK(near.init(p, pu, nv, repul_dist, r_near));
    V3D_CHECK(frep.alloc(3 * nv * ng));
    V3D_CHECK(vup.alloc(3 * nv * mu));
    V3D_CHECK(vpole.alloc(3 * nv * 2));
  }
  return work.alloc(std::max({shc2grid_work(3 * nv, p, pu), grid2shc_work(3 * nv, p, p),
                              shc2grid_work(3 * nv, p, p),
                              ord2 ? shc2grid_work(3 * nv, p, ord2) : 0L}));
}

int StokesG::setup(const double *X, hipStream_t st) {
  const long ng = sh_ngrid(p), mu = sh_ngrid(pu);
  near_off_ready = false;  // new sources
  V3D_CHECK(sht_grid2shc<double>(X, 3 * nv, p, p, shc, work, st));
  V3D_CHECK(sht_shc2grid<double>(shc, 3 * nv, p, pu, tmpx, tmpt, tmpp, work, st));
  V3D_CHECK(sht_shc2pole<double>(shc, 3 * nv, p, pole, st));
  if (!dqw.p) {
    std::vector<double> qw;
    legendre_weights(pu, qw);
    V3D_CHECK(upload(dqw, qw));
  }
  k_far_setup<<<gblocks(nv * (sh_ngrid(pu) + 2)), 256, 0, st>>>(
      tmpx, tmpt, tmpp, pole, dqw, nv, pu, far_src, aqw, want_dl ? nrm_up.p : nullptr);
  V3D_CHECK(launch_status());
  // targets: the re-synthesised p-grid positions (src/StokesVelocity.cc:227-239)
  const bool coarse = far_coarse > 0;
  V3D_CHECK(sht_shc2grid<double>(shc, 3 * nv, p, p, tmpx, coarse ? tmpt.p : nullptr,
                                 coarse ? tmpp.p : nullptr, work, st));
  k_to_points<<<gblocks(nv * ng), 256, 0, st>>>(tmpx, nv, ng, trg);
  V3D_CHECK(launch_status());
  if (coarse) {  // the targets double as the coarse far-field sources
    if (!dqwp.p) {
      std::vector<double> qwp;
      legendre_weights(p, qwp);
      V3D_CHECK(upload(dqwp, qwp));
    }
    k_coarse_weights<<<gblocks(nv * ng), 256, 0, st>>>(tmpt, tmpp, dqwp, nv, p, aqwp);
    V3D_CHECK(launch_status());
    long ng2 = 0;
    if (far_coarse2 > 0) {
      // the area element per unit-sphere measure, |x_t x x_p| / sin(theta),
      // and the degree-p surface at the order-q2 grid points
      if (!dinvs.p) {
        std::vector<double> x, w, invs(p + 1);
        gauss_legendre(p + 1, x, w);
        for (int i = 0; i <= p; ++i) invs[i] = 1.0 / std::sqrt(1.0 - x[i] * x[i]);
        V3D_CHECK(upload(dinvs, invs));
      }
      k_coarse_weights<<<gblocks(nv * ng), 256, 0, st>>>(tmpt, tmpp, dinvs, nv, p, c2gw);
      V3D_CHECK(launch_status());
      ng2 = sh_ngrid(ord2);
      V3D_CHECK(sht_shc2grid<double>(shc, 3 * nv, p, ord2, c2grid, nullptr, nullptr, work, st));
      k_to_points<<<gblocks(nv * ng2), 256, 0, st>>>(c2grid, nv, ng2, c2pos);
      V3D_CHECK(launch_status());
    }
    V3D_CHECK(far_coarse_flags(far_src, nv * (mu + 2), mu + 2, trg, nv * ng, ng, far_coarse,
                               far_coarse2, cflags, ccounts, st));
    unsigned long long c[3];
    V3D_HIP(hipMemcpyAsync(c, ccounts, sizeof(c), hipMemcpyDeviceToHost, st));
    V3D_HIP(hipStreamSynchronize(st));
    far_pairs = 64.0 * ((double)c[0] * ng + (double)c[1] * (mu + 2) + (double)c[2] * ng2);
  } else {
    // every target against every other vesicle's upsampled sources
    far_pairs = (double)(nv * ng) * (double)(nv - 1) * (double)(mu + 2);
  }
  if (use_near) V3D_CHECK(near.setup(far_src, trg, st));
  // the singular operators on bst, after everything st has queued (the
  // positions, and the last step's applications of the operators)
  V3D_HIP(hipEventRecord(ev_in, st));
  V3D_HIP(hipStreamWaitEvent(bst, ev_in, 0));
  if (precision != 1)
    V3D_CHECK(singular_build<double>(X, nv, p, pu, SL.p, want_dl ? DL.p : nullptr, bst));
  else
    V3D_CHECK(singular_build<float>(X, nv, p, pu, SL32, want_dl ? DL32 : nullptr, bst));
  if (precision == 2) {
    const long nmat = nv * 9 * (long)sh_ncoef(p) * sh_ncoef(p);
    k_round_f32<<<gblocks(nmat), 256, 0, bst>>>(SL.p, nmat, SL32);
    V3D_CHECK(launch_status());
  }
  V3D_HIP(hipEventRecord(ev_built, bst));
  build_pending = true;
  return hip_status(hipStreamSynchronize(st));
}

int StokesG::wait_build(hipStream_t st) {
  if (!build_pending) return VES3D_SUCCESS;
  V3D_HIP(hipStreamWaitEvent(st, ev_built, 0));
  build_pending = false;
  return VES3D_SUCCESS;
}

int StokesG::pack_coarse(const double *f, double *qcoarse, hipStream_t st) {
  const long ng = sh_ngrid(p);
  k_pack_coarse<<<gblocks(nv * ng), 256, 0, st>>>(f, aqwp, nv, ng, qcoarse);
  return launch_status();
}

// Level-2 densities: (f c2gw) truncated to degree ord2, synthesised at the
// order-ord2 grid points, times the order-ord2 quadrature weights.
int StokesG::pack_level2(const double *f, double *out, hipStream_t st) {
  const long ng = sh_ngrid(p), nc = sh_ncoef(p), ng2 = sh_ngrid(ord2);
  k_mul_pointwise<<<gblocks(3 * nv * ng), 256, 0, st>>>(f, c2gw, nv, ng, c2prod);
  V3D_CHECK(launch_status());
  V3D_CHECK(sht_grid2shc<double>(c2prod, 3 * nv, p, p, c2coef, work, st));
  k_mask_coef<<<gblocks(3 * nv * nc), 256, 0, st>>>(c2coef, c2mask, 3 * nv * nc, nc);
  V3D_CHECK(launch_status());
  V3D_CHECK(sht_shc2grid<double>(c2coef, 3 * nv, p, ord2, c2grid, nullptr, nullptr, work, st));
  k_pack_level2<<<gblocks(nv * ng2), 256, 0, st>>>(c2grid, c2w, nv, ng2, ord2, out);
  return launch_status();
}

FarCoarse StokesG::coarse_levels(const double *f1, const double *f2, hipStream_t st, int *err) {
  const long ng = sh_ngrid(p);
  FarCoarse fc{trg, qc, nullptr, (size_t)ng, cflags};
  *err = VES3D_SUCCESS;
  auto chk = [err](int e) {
    if (e != VES3D_SUCCESS && *err == VES3D_SUCCESS) *err = e;
  };
  if (f2 && !qc2.p) chk(qc2.alloc(nv * ng * 3));
  chk(pack_coarse(f1, qc, st));
  if (f2) {
    fc.den2 = qc2;
    chk(pack_coarse(f2, qc2, st));
  }
  if (far_coarse2 > 0) {
    const long ng2 = sh_ngrid(ord2);
    fc.pos2 = c2pos;
    fc.cpg2 = (size_t)ng2;
    fc.l2den1 = c2den;
    chk(pack_level2(f1, c2den, st));
    if (f2) {
      if (!c2den2.p) chk(c2den2.alloc(nv * ng2 * 3));
      fc.l2den2 = c2den2;
      chk(pack_level2(f2, c2den2, st));
    }
  }
  return fc;
}

int StokesG::density_coef(const double *f, double *Fc, hipStream_t st, const double *d) {
  if (f) V3D_CHECK(sht_grid2shc<double>(f, 3 * nv, p, p, Fc, work, st));
  if (d) V3D_CHECK(sht_grid2shc<double>(d, 3 * nv, p, p, G, work, st));
  return VES3D_SUCCESS;
}

int StokesG::self_velocity(bool sl, const double *Fc, double *Vc, double *vel, bool pack,
                           hipStream_t st, bool dl) {
  V3D_CHECK(wait_build(st));
  // the operators come from singular_build, whose (a, b) and (b, a) blocks
  // are equal: the apply reads the six blocks a <= b (sym)
  // precision 2, the self-interaction alone (an inner solve's operator, a
  // preconditioner's): the FP32 copy halves the HBM-bound apply
  if (precision == 2 && !pack && !dl)
    V3D_CHECK(self_apply<float>(sl ? SL32 : nullptr, Fc, nullptr, G.p, nv, p, Vc, st, true));
  else if (precision != 1)
    V3D_CHECK(self_apply<double>(sl ? SL.p : nullptr, Fc, dl ? DL.p : nullptr, G.p, nv, p, Vc,
                                 st, true));
  else
    V3D_CHECK(self_apply<float>(sl ? SL32 : nullptr, Fc, dl ? DL32 : nullptr, G.p, nv, p, Vc,
                                st, true));
  return vel ? sht_shc2grid<double>(Vc, 3 * nv, p, p, vel, nullptr, nullptr, work, st)
             : VES3D_SUCCESS;
}

int StokesG::self_coef(const double *f, double *Vc, hipStream_t st) {
  V3D_CHECK(density_coef(f, F, st, nullptr));
  return self_velocity(true, F, Vc, nullptr, false, st, false);
}

int StokesG::pack_density(bool sl, const double *Fc, double *upc, double *qc, hipStream_t st,
                          bool dl) {
  const long mu = sh_ngrid(pu);
  if (sl) {
    V3D_CHECK(sht_shc2grid<double>(Fc, 3 * nv, p, pu, upc, nullptr, nullptr, work, st));
    k_pack_density<<<gblocks(nv * (mu + 2)), 256, 0, st>>>(upc, aqw, nv, mu, qc);
    V3D_CHECK(launch_status());
  }
  if (dl) {  // uforce_double / qforce_double (src/StokesVelocity.cc:338-378)
    V3D_CHECK(sht_shc2grid<double>(G, 3 * nv, p, pu, dup, nullptr, nullptr, work, st));
    V3D_CHECK(sht_shc2pole<double>(G, 3 * nv, p, dpole, st));
    k_pack_dl<<<gblocks(nv * (mu + 2)), 256, 0, st>>>(dup, aqw, nrm_up, nv, mu, qd);
    V3D_CHECK(launch_status());
  }
  return VES3D_SUCCESS;
}

int StokesG::self_and_density(const double *f, double *Fc, double *Vc, double *vel, double *upc,
                              double *qc, bool pack, hipStream_t st, const double *d) {
  V3D_CHECK(density_coef(f, Fc, st, d));
  V3D_CHECK(self_velocity(f != nullptr, Fc, Vc, vel, pack, st, d != nullptr));
  if (!pack) return VES3D_SUCCESS;
  return pack_density(f != nullptr, Fc, upc, qc, st, d != nullptr);
}

int StokesG::near_correct(const double *qc, const double *Vc, double *vel, hipStream_t st,
                          bool dl) {
  // near-singular correction with the self velocity on the upsampled grid
  // and at the poles (S_vel_up, src/StokesVelocity.cc:452-470)
  V3D_CHECK(sht_shc2grid<double>(Vc, 3 * nv, p, pu, vup, nullptr, nullptr, work, st));
  V3D_CHECK(sht_shc2pole<double>(Vc, 3 * nv, p, vpole, st));
  return near.apply(qc, dl ? qd.p : nullptr, vup, vpole, dl ? dup.p : nullptr,
                    dl ? dpole.p : nullptr, vel, st);
}

int StokesG::enable_dl() {
  if (want_dl) return VES3D_SUCCESS;
  const long nc = sh_ncoef(p), mu = sh_ngrid(pu), ng = sh_ngrid(p);
  const long nmat = nv * 9 * nc * nc;
  if (precision != 1) {
    V3D_CHECK(DL.alloc(nmat));
  } else {
    if (hipMalloc(&DL32, sizeof(float) * nmat) != hipSuccess) return VES3D_MEMORY_ERROR;
  }
  V3D_CHECK(nrm_up.alloc(3 * nv * mu));
  V3D_CHECK(G.alloc(3 * nv * nc));
  V3D_CHECK(qd.alloc(nv * (mu + 2) * 6));
  V3D_CHECK(dup.alloc(3 * nv * mu));
  V3D_CHECK(dpole.alloc(3 * nv * 2));
  V3D_CHECK(selfv.alloc(3 * nv * ng));
  if (!vup.p) {
    V3D_CHECK(vup.alloc(3 * nv * mu));
    V3D_CHECK(vpole.alloc(3 * nv * 2));
  }
  want_dl = true;
  return VES3D_SUCCESS;
}

int StokesG::eval_off(const double *f, const double *d, const double *trgo, long n, double *vel,
                      hipStream_t st) {
  const long mu = sh_ngrid(pu);
  if ((!f && !d) || (d && !want_dl) || !trgo || !vel || n < 0) return VES3D_INVALID_PARAMETER;
  if (n == 0) return VES3D_SUCCESS;
  if (!selfv.p) V3D_CHECK(selfv.alloc(3 * nv * sh_ngrid(p)));
  if (!vup.p) {
    V3D_CHECK(vup.alloc(3 * nv * mu));
    V3D_CHECK(vpole.alloc(3 * nv * 2));
  }
  V3D_CHECK(self_and_density(f, F, V, selfv, up, q, true, st, d));
  V3D_CHECK(far_field_sl_dl(false, far_src, f ? q.p : nullptr, d ? qd.p : nullptr, nv * (mu + 2),
                            trgo, vel, (size_t)n, 0, 0, true, st));
  // near-singular correction (near_singular1): the same r_near as the
  // surface targets'
  if (!near_off_ready) {
    const double rn = use_near ? near.r_near : 0.0;
    V3D_CHECK(near_off.init(p, pu, nv, 0.0, rn, false, 256));
    near_off_ready = true;
  }
  V3D_CHECK(near_off.setup(far_src, trgo, st, n));
  if (near_off.npairs == 0) return VES3D_SUCCESS;
  V3D_CHECK(sht_shc2grid<double>(V, 3 * nv, p, pu, vup, nullptr, nullptr, work, st));
  V3D_CHECK(sht_shc2pole<double>(V, 3 * nv, p, vpole, st));
  return near_off.apply(f ? q.p : nullptr, d ? qd.p : nullptr, vup, vpole, d ? dup.p : nullptr,
                        d ? dpole.p : nullptr, vel, st);
}

int StokesG::event_begin(int kind, hipStream_t st) {
  if (!timing) return VES3D_SUCCESS;
  if (nev == (int)ev0.size()) {
    hipEvent_t a, b;
    V3D_HIP(hipEventCreate(&a));
    V3D_HIP(hipEventCreate(&b));
    ev0.push_back(a);
    ev1.push_back(b);
    evkind.push_back(0);
  }
  evkind[nev] = kind;
  V3D_HIP(hipEventRecord(ev0[nev], st));
  return VES3D_SUCCESS;
}
int StokesG::event_end(hipStream_t st) {
  if (!timing) return VES3D_SUCCESS;
  V3D_HIP(hipEventRecord(ev1[nev++], st));
  return VES3D_SUCCESS;
}

int StokesG::apply(const double *f, double *vel, hipStream_t st, bool add_repul, bool local,
                   const double *d) {
  const long ng = sh_ngrid(p), mu = sh_ngrid(pu);
  const bool near_on = use_near && near.npairs > 0;
  if ((d && !want_dl) || (!f && (!d || add_repul))) return VES3D_INVALID_PARAMETER;
  if (add_repul && near_on && near.repul > 0) {
    V3D_HIP(hipMemcpyAsync(frep, f, sizeof(double) * 3 * nv * ng, hipMemcpyDeviceToDevice, st));
    V3D_CHECK(near.add_repulsion(frep, st));
    f = frep;
  }
  V3D_CHECK(self_and_density(f, F, V, vel, up, q, !local, st, d));
  if (local) return VES3D_SUCCESS;
  if (d) {
    // single and double layer over the upsampled sources (qforce_single,
    // qforce_double), own vesicle skipped
    V3D_CHECK(event_begin(1, st));
    V3D_CHECK(far_field_sl_dl(precision != 0, far_src, f ? q.p : nullptr, qd, nv * (mu + 2), trg,
                              farv, nv * ng, mu + 2, ng, false, st));
    V3D_CHECK(event_end(st));
    k_add_points<<<gblocks(nv * ng), 256, 0, st>>>(farv, nv, ng, vel);
    V3D_CHECK(launch_status());
    if (!near_on) return VES3D_SUCCESS;
    V3D_CHECK(event_begin(3, st));
    V3D_CHECK(near_correct(f ? q.p : nullptr, V, vel, st, true));
    return event_end(st);
  }
  FarCoarse fc{};
  if (far_coarse > 0) {
    int e = VES3D_SUCCESS;
    fc = coarse_levels(f, nullptr, st, &e);
    V3D_CHECK(e);
  }
  V3D_CHECK(event_begin(1, st));
  V3D_CHECK(far_field_between_groups(precision != 0, far_src, q, nv * (mu + 2), trg, farv, nv * ng,
                                     mu + 2, ng, st, far_coarse > 0 ? &fc : nullptr));
  V3D_CHECK(event_end(st));
  k_add_points<<<gblocks(nv * ng), 256, 0, st>>>(farv, nv, ng, vel);
  V3D_CHECK(launch_status());
  if (!near_on) return VES3D_SUCCESS;
  V3D_CHECK(event_begin(3, st));
  V3D_CHECK(near_correct(q, V, vel, st));
  return event_end(st);
}

int StokesG::apply2(const double *f1, double *vel1, bool add_repul1, const double *f2,
                    double *vel2, hipStream_t st) {
  const long ng = sh_ngrid(p), mu = sh_ngrid(pu);
  const bool near_on = use_near && near.npairs > 0;
  if (!F2.p) {
    for (DBuf *b : {&F2, &V2}) V3D_CHECK(b->alloc(3 * nv * sh_ncoef(p)));
    V3D_CHECK(up2.alloc(3 * nv * mu));
    V3D_CHECK(q2.alloc(nv * (mu + 2) * 3));
    V3D_CHECK(farv2.alloc(nv * ng * 3));
  }
  if (add_repul1 && near_on && near.repul > 0) {
    V3D_HIP(hipMemcpyAsync(frep, f1, sizeof(double) * 3 * nv * ng, hipMemcpyDeviceToDevice, st));
    V3D_CHECK(near.add_repulsion(frep, st));
    f1 = frep;
  }
  // the far-field pass first: it needs no self interaction, so it runs
  // while the singular build of the last setup finishes on bst
  V3D_CHECK(density_coef(f1, F, st, nullptr));
  V3D_CHECK(density_coef(f2, F2, st, nullptr));
  V3D_CHECK(pack_density(true, F, up, q, st, false));
  V3D_CHECK(pack_density(true, F2, up2, q2, st, false));
  FarCoarse fc{};
  if (far_coarse > 0) {
    int e = VES3D_SUCCESS;
    fc = coarse_levels(f1, f2, st, &e);
    V3D_CHECK(e);
  }
  V3D_CHECK(event_begin(2, st));
  V3D_CHECK(far_field_between_groups2(precision != 0, far_src, q, q2, nv * (mu + 2), trg, farv,
                                      farv2, nv * ng, mu + 2, ng, st,
                                      far_coarse > 0 ? &fc : nullptr));
  V3D_CHECK(event_end(st));
  V3D_CHECK(self_velocity(true, F, V, vel1, true, st, false));
  V3D_CHECK(self_velocity(true, F2, V2, vel2, true, st, false));
  k_add_points<<<gblocks(nv * ng), 256, 0, st>>>(farv, nv, ng, vel1);
  V3D_CHECK(launch_status());
  k_add_points<<<gblocks(nv * ng), 256, 0, st>>>(farv2, nv, ng, vel2);
  V3D_CHECK(launch_status());
  if (!near_on) return VES3D_SUCCESS;
  V3D_CHECK(event_begin(3, st));
  V3D_CHECK(near_correct(q, V, vel1, st));
  V3D_CHECK(near_correct(q2, V2, vel2, st));
  return event_end(st);
}

}  // namespace v3d

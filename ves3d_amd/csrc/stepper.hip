// stepper.hip -- the globally implicit interfacial-velocity time step
// (InterfacialVelocity::updateImplicit with solve_for_velocity,
// src/InterfacialVelocity.cc:203-560) on one GPU, and the BiCGStab solver
// (src/BiCGStab.cc:45-138) with apply / precondition callbacks shaped like
// ParallelLinOp::apply_type and ParallelLinSolver::precond_type
// (include/ParallelLinSolverInterface.h:111, :160) but on device vectors.
//
// Step: geometry at p and 2p, singular operators and far-field setup,
// right-hand side  b = P[u_inf + S F_b(x) ; div(.)],  solve  A (u, sigma) = b
// with the Galerkin matvec  A = P [u - S f(dt u, sigma) ; -div S f] P^-1,
// then x += dt u.  P packs grid fields into SH coefficients in the
// reference's order-major Galerkin layout (src/InterfacialVelocity.cc:981-1052).
#include <algorithm>
#include <chrono>
#include <cmath>
#include <vector>

#include "implicit.hpp"
#include "learned_pc.hpp"
#include "runtime.hpp"
#include "sh_tables.hpp"
#include "sht.hpp"
#include "vecops.hpp"

namespace v3d {

namespace {
inline unsigned gb(long n) {
  long b = (n + 255) / 256;
  if (b > 65536) b = 65536;
  return (unsigned)(b < 1 ? 1 : b);
}

// Canonical frame of every vesicle for the learned preconditioner: the
// orthogonal polar factor Q of C = int X (x) s dA_param, s = (sin t cos p,
// sin t sin p, cos t) the parameter sphere (its degree-1 harmonics) and
// fgrid[4][ng] = (s_x, s_y, s_z, Gauss-Legendre weight) per grid point.  A
// rotated copy R X of a shape has C' = R C and frame R Q, so F^T x is the
// same for both.  One workgroup per vesicle; C is reduced in LDS and thread 0
// runs the scaled Newton iteration Q <- (g Q + Q^-T / g) / 2 (Higham 1986,
// g = sqrt(|Q^-1|_F / |Q|_F)), which converges quadratically for the
// nonsingular C of a closed surface; a singular C gives the identity.
__device__ inline double inv3(const double *a, double *r) {  // r = a^-1, returns det
  r[0] = a[4] * a[8] - a[5] * a[7];
  r[1] = a[2] * a[7] - a[1] * a[8];
  r[2] = a[1] * a[5] - a[2] * a[4];
  r[3] = a[5] * a[6] - a[3] * a[8];
  r[4] = a[0] * a[8] - a[2] * a[6];
  r[5] = a[2] * a[3] - a[0] * a[5];
  r[6] = a[3] * a[7] - a[4] * a[6];
  r[7] = a[1] * a[6] - a[0] * a[7];
  r[8] = a[0] * a[4] - a[1] * a[3];
  const double det = a[0] * r[0] + a[1] * r[3] + a[2] * r[6];
  const double id = det != 0.0 ? 1.0 / det : 0.0;
  for (int k = 0; k < 9; ++k) r[k] *= id;
  return det;
}
__global__ __launch_bounds__(256) void k_frames(const double *__restrict__ X, long ng,
                                                const double *__restrict__ fgrid,
                                                double *__restrict__ frames) {
  const long v = blockIdx.x;
  const double *Xv = X + v * 3 * ng;
  double acc[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  for (long j = threadIdx.x; j < ng; j += 256) {
    const double w = fgrid[3 * ng + j];
    const double s0 = fgrid[j] * w, s1 = fgrid[ng + j] * w, s2 = fgrid[2 * ng + j] * w;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      const double x = Xv[i * ng + j];
      acc[3 * i] += x * s0;
      acc[3 * i + 1] += x * s1;
      acc[3 * i + 2] += x * s2;
    }
  }
  __shared__ double red[4][9];
#pragma unroll
  for (int k = 0; k < 9; ++k) {
    const double t = wave_sum(acc[k]);
    if (threadIdx.x % 64 == 0) red[threadIdx.x / 64][k] = t;
  }
  __syncthreads();
  if (threadIdx.x != 0) return;
  double q[9], qi[9];
  double nq = 0;
  for (int k = 0; k < 9; ++k) {
    q[k] = red[0][k] + red[1][k] + red[2][k] + red[3][k];
    nq += q[k] * q[k];
  }
  bool ok = nq > 0;
  for (int it = 0; ok && it < 100; ++it) {
    const double det = inv3(q, qi);
    if (!(fabs(det) > 1e-12 * nq * sqrt(nq))) {
      ok = false;
      break;
    }
    double a = 0, b = 0;
    for (int k = 0; k < 9; ++k) {
      a += q[k] * q[k];
      b += qi[k] * qi[k];
    }
    const double g = sqrt(sqrt(b / a));
    double diff = 0;
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) {
        const double nw = 0.5 * (g * q[3 * r + c] + qi[3 * c + r] / g);
        diff += (nw - q[3 * r + c]) * (nw - q[3 * r + c]);
        q[3 * r + c] = nw;
      }
    nq = 3.0;
    if (diff < 1e-30) break;
  }
  for (int k = 0; k < 9; ++k) frames[9 * v + k] = ok ? q[k] : (k % 4 == 0 ? 1.0 : 0.0);
}

// Function-major coefficients [nf][nc] <-> order-major Galerkin layout.
// cinfo[c] = (block offset, block length) of coefficient c.
__global__ void k_fm_to_om(const double *__restrict__ fm, long nf, int nc,
                           const int2 *__restrict__ cinfo, double *__restrict__ om) {
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < nf * nc;
       i += (long)gridDim.x * blockDim.x) {
    const long f = i / nc;
    const int c = (int)(i % nc);
    const int2 bl = cinfo[c];
    om[nf * bl.x + f * bl.y + (c - bl.x)] = fm[i];
  }
}
__global__ void k_om_to_fm(const double *__restrict__ om, long nf, int nc,
                           const int2 *__restrict__ cinfo, double *__restrict__ fm) {
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < nf * nc;
       i += (long)gridDim.x * blockDim.x) {
    const long f = i / nc;
    const int c = (int)(i % nc);
    const int2 bl = cinfo[c];
    fm[i] = om[nf * bl.x + f * bl.y + (c - bl.x)];
  }
}
// ---- after the solve: reparametrization and area / volume correction
// S[f][c] *= w[c] (function-major coefficients, w per coefficient).
__global__ void k_scale_fm(double *S, long nf, int nc, const double *__restrict__ w) {
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < nf * nc;
       i += (long)gridDim.x * blockDim.x)
    S[i] *= w[i % nc];
}
// u -= (u . n) n on AxisMajor [nv][3][ng] fields.
__global__ void k_tangent(double *u, const double *__restrict__ n, long nv, long ng) {
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < nv * ng;
       i += (long)gridDim.x * blockDim.x) {
    const long v = i / ng, g = i % ng, b = v * 3 * ng + g;
    const double d = u[b] * n[b] + u[b + ng] * n[b + ng] + u[b + 2 * ng] * n[b + 2 * ng];
    u[b] -= d * n[b];
    u[b + ng] -= d * n[b + ng];
    u[b + 2 * ng] -= d * n[b + 2 * ng];
  }
}
// Per vesicle (one workgroup each): u /= max_g |u_g|, then the degree-weighted
// coefficient inner products xx, xu, uu of x and u (Xc, Uc: function-major
// coefficients of the 3 components; the scaling of u enters uu and xu).
__global__ __launch_bounds__(256) void k_rep_norm_inner(double *u, long ng,
                                                        const double *__restrict__ Xc,
                                                        double *Uc, int nc,
                                                        const double *__restrict__ w,
                                                        double *out3) {
  const long v = blockIdx.x;
  double *uv = u + v * 3 * ng;
  __shared__ double sh[3][4];
  __shared__ double smax;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  double m = 0;
  for (long g = threadIdx.x; g < ng; g += 256) {
    const double a = uv[g], b = uv[ng + g], c = uv[2 * ng + g];
    m = fmax(m, sqrt(a * a + b * b + c * c));
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o, 64));
  if (lane == 0) sh[0][wv] = m;
  __syncthreads();
  if (threadIdx.x == 0) smax = fmax(fmax(sh[0][0], sh[0][1]), fmax(sh[0][2], sh[0][3]));
  __syncthreads();
  const double inv = 1.0 / smax;
  for (long g = threadIdx.x; g < 3 * ng; g += 256) uv[g] *= inv;
  const double *xc = Xc + v * 3 * nc;
  double *uc = Uc + v * 3 * nc;
  double xx = 0, xu = 0, uu = 0;
  for (long c = threadIdx.x; c < 3L * nc; c += 256) {
    const double ww = w[c % nc], a = xc[c], b = uc[c] * inv;
    uc[c] = b;
    xx += ww * a * a;
    xu += ww * a * b;
    uu += ww * b * b;
  }
  __syncthreads();
  xx = wave_sum(xx);
  xu = wave_sum(xu);
  uu = wave_sum(uu);
  if (lane == 0) {
    sh[0][wv] = xx;
    sh[1][wv] = xu;
    sh[2][wv] = uu;
  }
  __syncthreads();
  if (threadIdx.x < 3) {
    const int k = threadIdx.x;
    out3[v * 3 + k] = ((sh[k][0] + sh[k][1]) + sh[k][2]) + sh[k][3];
  }
}
// Per-vesicle step of the reparametrization (src/InterfacialVelocity.cc:
// 1728-1741): dt = min(ts, -xu/uu), zero (and the vesicle frozen) when below
// tol or frozen before (E == 0 with have_E).
__global__ void k_rep_step(const double *__restrict__ in3, long nv, double ts, double tol,
                           int have_E, double *E, double *dt) {
  for (long v = (long)blockIdx.x * blockDim.x + threadIdx.x; v < nv;
       v += (long)gridDim.x * blockDim.x) {
    double xx = in3[3 * v], d = fmin(ts, -in3[3 * v + 1] / in3[3 * v + 2]);
    if (d < tol || (have_E && E[v] == 0.0)) {
      xx = 0.0;
      d = 0.0;
    }
    dt[v] = d;
    E[nv + v] = xx;  // candidate energy, committed by the host if the loop goes on
  }
}
// x += dt[v] u.
__global__ void k_rep_update(double *x, const double *__restrict__ u,
                             const double *__restrict__ dt, long nv, long n3) {
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < nv * n3;
       i += (long)gridDim.x * blockDim.x)
    x[i] += dt[i / n3] * u[i];
}
// Per vesicle on the upsampled surface: area, volume and the integrals of
// (-2H)^2, -2H and 1 against the area element (GaussLegendreIntegrator,
// src/EvolveSurface.cc:352-440): out4 = [A, V, h2, h1].
__global__ __launch_bounds__(256) void k_av_integrals(const double *__restrict__ X,
                                                      const double *__restrict__ nrm,
                                                      const double *__restrict__ W,
                                                      const double *__restrict__ H,
                                                      const double *__restrict__ qw, long ng,
                                                      int nlon, double *out4) {
  const long v = blockIdx.x;
  __shared__ double sh[4][4];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  double a = 0, vol = 0, h2 = 0, h1 = 0;
  for (long g = threadIdx.x; g < ng; g += 256) {
    const long b = v * 3 * ng + g;
    const double dw = W[v * ng + g] * qw[g / nlon];
    const double xn = X[b] * nrm[b] + X[b + ng] * nrm[b + ng] + X[b + 2 * ng] * nrm[b + 2 * ng];
    const double hx = -2.0 * H[v * ng + g];
    a += dw;
    vol += xn / 3.0 * dw;
    h2 += hx * hx * dw;
    h1 += hx * dw;
  }
  a = wave_sum(a);
  vol = wave_sum(vol);
  h2 = wave_sum(h2);
  h1 = wave_sum(h1);
  if (lane == 0) {
    sh[0][wv] = a;
    sh[1][wv] = vol;
    sh[2][wv] = h2;
    sh[3][wv] = h1;
  }
  __syncthreads();
  if (threadIdx.x < 4) {
    const int k = threadIdx.x;
    out4[v * 4 + k] = ((sh[k][0] + sh[k][1]) + sh[k][2]) + sh[k][3];
  }
}
// Newton update of the area / volume correction: per vesicle
// [dA/dX dA/dY; dV/dX dV/dY] (dX, dY) = (area err, vol err), then
// x_up += (dX (-2H) + dY) n.
__global__ void k_av_update(double *Xu, const double *__restrict__ nrm,
                            const double *__restrict__ H, const double *__restrict__ in4,
                            const double *__restrict__ a0, const double *__restrict__ v0,
                            long nv, long ng) {
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < nv * ng;
       i += (long)gridDim.x * blockDim.x) {
    const long v = i / ng, g = i % ng;
    const double *q = in4 + 4 * v;
    const double dAdX = q[2], dAdY = q[3], dVdX = q[3], dVdY = q[0];
    const double ae = a0[v] - q[0], ve = v0[v] - q[1];
    const double di = 1.0 / (dAdX * dVdY - dAdY * dVdX);
    const double dX = dVdY * di * ae + (-dAdY * di) * ve;
    const double dY = (-dVdX * di) * ae + dAdX * di * ve;
    const double s = dX * (-2.0 * H[v * ng + g]) + dY;
    const long b = v * 3 * ng + g;
    Xu[b] += s * nrm[b];
    Xu[b + ng] += s * nrm[b + ng];
    Xu[b + 2 * ng] += s * nrm[b + 2 * ng];
  }
}

// ScaleFreq on the order-major layout (src/SHTrans.cc:192-223).
__global__ void k_scale_om(const double *__restrict__ x, long nf, int nc,
                           const int2 *__restrict__ cinfo, const double *__restrict__ scale,
                           double *__restrict__ y) {
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < nf * nc;
       i += (long)gridDim.x * blockDim.x) {
    const long f = i / nc;
    const int c = (int)(i % nc);
    const int2 bl = cinfo[c];
    const long o = nf * bl.x + f * bl.y + (c - bl.x);
    y[o] = x[o] * scale[c];
  }
}
// out[v][:] = c[v] x[v][:] on vector fields [nv][3 ng] (per-vesicle factors)
__global__ void k_ves_scale(const double *__restrict__ x, const double *__restrict__ c, long nv,
                            long n3, double *__restrict__ out) {
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < nv * n3;
       i += (long)gridDim.x * blockDim.x)
    out[i] = c[i / n3] * x[i];
}
// y[v][:] = c[v] y[v][:] - s[v][:]
__global__ void k_ves_axmy(double *y, const double *__restrict__ c, const double *__restrict__ s,
                           long nv, long n3) {
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < nv * n3;
       i += (long)gridDim.x * blockDim.x)
    y[i] = c[i / n3] * y[i] - s[i];
}

// u_inf = (rate * z, 0, 0)  (src/BgFlow.cc:58-76)
__global__ void k_shear(const double *__restrict__ X, long nv, long ng, double rate,
                        double *__restrict__ u) {
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < nv * ng;
       i += (long)gridDim.x * blockDim.x) {
    const long v = i / ng, g = i % ng;
    u[(v * 3) * ng + g] = rate * X[(v * 3 + 2) * ng + g];
    u[(v * 3 + 1) * ng + g] = 0.0;
    u[(v * 3 + 2) * ng + g] = 0.0;
  }
}
}  // namespace

typedef int (*apply_fn)(const void *ctx, const double *x, double *y);

// BiCGStab (src/BiCGStab.cc:45-138), right-preconditioned, FP64 dots.
// Returns Error_t; *status gets the BiCGSReturn code (include/BiCGStab.h:20-24).
// work: 8 n doubles of caller-owned scratch (the step owns one for its
// solves), or null to allocate.  x_zero: x holds zeros, so the initial
// residual is b and the first operator application is skipped.
int bicgstab(long n, apply_fn A, const void *actx, apply_fn P, const void *pctx, double *x,
             const double *b, int *max_iter, double *tol, int *status, hipStream_t st,
             const double *Ax0 = nullptr, double *work = nullptr, bool x_zero = false) {
  DBuf own[8];
  double *vec[8];
  for (int i = 0; i < 8; ++i) {
    if (work) {
      vec[i] = work + i * n;
    } else {
      V3D_CHECK(own[i].alloc(n));
      vec[i] = own[i].p;
    }
  }
  double *r = vec[0], *rt = vec[1], *p = vec[2], *v = vec[3], *s = vec[4], *t = vec[5],
         *ph = vec[6], *sh = vec[7];
  int err = 0;
  double normb = dev_norm(n, b, st, &err);
  V3D_CHECK(err);
  if (normb == 0.0) normb = 1.0;
  if (x_zero) {
    V3D_CHECK(dev_copy(n, b, r, st));  // r = b - A 0
  } else {
    if (Ax0)  // A x computed by the caller (together with b)
      V3D_CHECK(dev_copy(n, Ax0, r, st));
    else
      V3D_CHECK(A(actx, x, r));
    V3D_CHECK(dev_axpby(n, 1.0, b, -1.0, r, st));  // r = b - Ax
  }
  V3D_CHECK(dev_copy(n, r, rt, st));
  double resid = dev_norm(n, r, st, &err) / normb;
  V3D_CHECK(err);
  if (resid <= *tol) {
    *tol = resid;
    *max_iter = 0;
    *status = 100;
    return VES3D_SUCCESS;
  }
  double rho1 = 0, rho2 = 1, alpha = 1, omega = 1, beta = 0;
  const int maxit = *max_iter;
  for (int i = 1; i <= maxit; ++i) {
    if (resid != resid) {
      *status = 104;
      *tol = resid;
      *max_iter = i;
      return VES3D_SUCCESS;
    }
    {
      const double *xs[1] = {rt}, *ys[1] = {r};
      V3D_CHECK(dev_dots(n, 1, xs, ys, &rho1, st));
    }
    if (rho1 == 0) {
      *tol = dev_norm(n, r, st, &err) / normb;
      *max_iter = i;
      *status = 102;
      return err;
    }
    if (i == 1) {
      V3D_CHECK(dev_copy(n, r, p, st));
    } else {
      beta = (rho1 / rho2) * (alpha / omega);
      V3D_CHECK(dev_axpby(n, -omega, v, 1.0, p, st));  // p = p - omega v
      V3D_CHECK(dev_axpby(n, 1.0, r, beta, p, st));    // p = r + beta p
    }
    if (P) V3D_CHECK(P(pctx, p, ph));
    else V3D_CHECK(dev_copy(n, p, ph, st));
    V3D_CHECK(A(actx, ph, v));
    double rtv = 0;
    {
      const double *xs[1] = {rt}, *ys[1] = {v};
      V3D_CHECK(dev_dots(n, 1, xs, ys, &rtv, st));
    }
    alpha = rho1 / rtv;
    V3D_CHECK(dev_axpy3(n, -alpha, v, r, s, st));  // s = r - alpha v
    resid = dev_norm(n, s, st, &err) / normb;
    V3D_CHECK(err);
    if (resid < *tol) {
      V3D_CHECK(dev_axpby(n, alpha, ph, 1.0, x, st));
      *tol = resid;
      *max_iter = i;
      *status = 100;
      return VES3D_SUCCESS;
    }
    if (P) V3D_CHECK(P(pctx, s, sh));
    else V3D_CHECK(dev_copy(n, s, sh, st));
    V3D_CHECK(A(actx, sh, t));
    double d2[2];
    {
      const double *xs[2] = {t, t}, *ys[2] = {s, t};
      V3D_CHECK(dev_dots(n, 2, xs, ys, d2, st));
    }
    omega = d2[0] / d2[1];
    V3D_CHECK(dev_axpby(n, alpha, ph, 1.0, x, st));
    V3D_CHECK(dev_axpby(n, omega, sh, 1.0, x, st));
    V3D_CHECK(dev_axpy3(n, -omega, t, s, r, st));  // r = s - omega t
    rho2 = rho1;
    resid = dev_norm(n, r, st, &err) / normb;
    V3D_CHECK(err);
    if (resid < *tol) {
      *tol = resid;
      *max_iter = i;
      *status = 100;
      return VES3D_SUCCESS;
    }
    if (omega == 0) {
      *tol = resid;
      *max_iter = i;
      *status = 103;
      return VES3D_SUCCESS;
    }
  }
  *tol = resid;
  *status = 101;
  return VES3D_SUCCESS;
}

// Flexible GMRES (Saad, SIAM J. Sci. Comput. 14 (1993) 461-469), right
// preconditioned, restarted every m iterations: the method of the
// reference's globally implicit solve (PETSc KSP GMRES,
// src/ParallelLinSolver_Petsc.cc:545-551, restart 1000, with a shell
// preconditioner), in the flexible form that a preconditioner which changes
// between applications (an inner solve, a network) needs: the preconditioned
// vectors z_j are kept and the iterate is x_0 + Z_k y.  The Arnoldi vectors
// are orthogonalised by classical Gram-Schmidt applied twice (batched FP64
// dots, fixed reduction order); Givens rotations give each iterate's
// residual norm |g_{j+1}|.  At the end of a cycle the residual vector is
// formed as V_{k+1} (beta e_1 - Hbar_k y), which equals b - A x_k up to
// rounding (the Arnoldi relation A Z_k = V_{k+1} Hbar_k holds for the
// operator applications made), so neither a restart nor the final residual
// costs an operator application; a cycle whose explicit residual misses the
// tolerance is followed by another.
// work: (2m + 1) n doubles (V: m + 1 vectors, Z: m), or null to allocate.
// *status: 100 converged, 101 iteration limit, 104 NaN (the BiCGSReturn
// codes, include/BiCGStab.h:20-24); *tol: the explicit residual ||b - A x||
// / ||b|| of the returned x; *max_iter: iterations (operator applications
// after the initial residual).
int fgmres(long n, apply_fn A, const void *actx, apply_fn P, const void *pctx, double *x,
           const double *b, int *max_iter, double *tol, int *status, hipStream_t st, int m,
           const double *Ax0 = nullptr, double *work = nullptr) {
  if (n < 1 || m < 1 || m >= kMaxVecCoefs) return VES3D_INVALID_PARAMETER;
  DBuf own;
  double *V = work;
  if (!V) {
    V3D_CHECK(own.alloc((2L * m + 1) * n));
    V = own.p;
  }
  double *Z = V + (long)(m + 1) * n;
  int err = 0;
  double normb = dev_norm(n, b, st, &err);
  V3D_CHECK(err);
  if (normb == 0.0) normb = 1.0;
  if (Ax0)  // A x computed by the caller (together with b)
    V3D_CHECK(dev_copy(n, Ax0, V, st));
  else
    V3D_CHECK(A(actx, x, V));
  V3D_CHECK(dev_axpby(n, 1.0, b, -1.0, V, st));  // V_0 = b - A x
  double beta = dev_norm(n, V, st, &err);
  V3D_CHECK(err);
  double resid = beta / normb;
  const double tol0 = *tol;
  const int maxit = *max_iter;
  int it = 0;
  // R: the Hessenberg matrix rotated to upper triangular; Hb: as built
  std::vector<double> R((m + 1) * m), Hb((m + 1) * m), cs(m), sn(m), g(m + 1), y(m),
      h(m + 1), h2(m + 1), c(m + 1);
  for (;;) {
    if (resid != resid) {
      *status = 104;
      break;
    }
    if (resid <= tol0) {
      *status = 100;
      break;
    }
    if (it >= maxit) {
      *status = 101;
      break;
    }
    V3D_CHECK(dev_axpby(n, 1.0 / beta, V, 0.0, V, st));  // v_0 = r / beta
    std::fill(g.begin(), g.end(), 0.0);
    g[0] = beta;
    int k = 0;
    while (k < m && it < maxit) {
      const double *vk = V + (long)k * n;
      double *zk = Z + (long)k * n, *w = V + (long)(k + 1) * n;
      if (P)
        V3D_CHECK(P(pctx, vk, zk));
      else
        V3D_CHECK(dev_copy(n, vk, zk, st));
      V3D_CHECK(A(actx, zk, w));
      ++it;
      // w -= V_k (V_k^T w), twice
      V3D_CHECK(dev_mdots(n, k + 1, V, n, w, h.data(), st));
      for (int i = 0; i <= k; ++i) c[i] = -h[i];
      V3D_CHECK(dev_combine(n, k + 1, V, n, c.data(), 1.0, w, st));
      V3D_CHECK(dev_mdots(n, k + 1, V, n, w, h2.data(), st));
      for (int i = 0; i <= k; ++i) {
        c[i] = -h2[i];
        h[i] += h2[i];
      }
      V3D_CHECK(dev_combine(n, k + 1, V, n, c.data(), 1.0, w, st));
      const double hn = dev_norm(n, w, st, &err);
      V3D_CHECK(err);
      h[k + 1] = hn;
      for (int i = 0; i <= k + 1; ++i) Hb[i * m + k] = h[i];
      for (int i = 0; i < k; ++i) {
        const double a = cs[i] * h[i] + sn[i] * h[i + 1];
        h[i + 1] = -sn[i] * h[i] + cs[i] * h[i + 1];
        h[i] = a;
      }
      const double den = std::hypot(h[k], h[k + 1]);
      cs[k] = den == 0.0 ? 1.0 : h[k] / den;
      sn[k] = den == 0.0 ? 0.0 : h[k + 1] / den;
      h[k] = den;
      g[k + 1] = -sn[k] * g[k];
      g[k] = cs[k] * g[k];
      for (int i = 0; i <= k; ++i) R[i * m + k] = h[i];
      ++k;
      if (hn > 0.0) V3D_CHECK(dev_axpby(n, 1.0 / hn, w, 0.0, w, st));  // v_k
      resid = std::fabs(g[k]) / normb;
      if (resid != resid || resid <= tol0 || hn == 0.0) break;
    }
    // y = R^-1 g, x += Z y
    for (int i = k - 1; i >= 0; --i) {
      double t = g[i];
      for (int j = i + 1; j < k; ++j) t -= R[i * m + j] * y[j];
      y[i] = R[i * m + i] != 0.0 ? t / R[i * m + i] : 0.0;
    }
    V3D_CHECK(dev_combine(n, k, Z, n, y.data(), 1.0, x, st));
    // r = V_{k+1} (beta e_1 - Hbar y), formed in Z_0, then moved to V_0
    for (int i = 0; i <= k; ++i) {
      double t = i == 0 ? beta : 0.0;
      for (int j = 0; j < k; ++j) t -= Hb[i * m + j] * y[j];
      c[i] = t;
    }
    V3D_CHECK(dev_combine(n, k + 1, V, n, c.data(), 0.0, Z, st));
    V3D_CHECK(dev_copy(n, Z, V, st));
    beta = dev_norm(n, V, st, &err);
    V3D_CHECK(err);
    resid = beta / normb;
  }
  *tol = resid;
  *max_iter = it;
  return VES3D_SUCCESS;
}

// ------------------------------------------------------------- the step
struct Stepper {
  ves3d_step_params prm;
  long nv = 0, ng = 0, nc = 0, n = 0;
  hipStream_t st = nullptr;
  ForcesG forces;
  StokesG stokes;
  DBuf X, U, sig;                  // state
  DBuf f, f2, Sf, ten, ug, sg, cf, rhs, xsol, yres, tmpD, yloc;
  DBuf lcu, lcd, lcv, lcs, lct;  // coefficients of the local operator (matvec_local_c)
  DBuf ub, fb, Sfb, Ax0;           // right-hand side fused with the first matvec
  DBuf bprec, tprec;               // diagonal-spectral preconditioner
  DBuf fgrid, frames;              // canonical frames of the vesicles (learned PC)
  int2 *cinfo = nullptr;
  LearnedPC *lpc = nullptr;        // not owned
  bool collect = false;            // record (A_loc e, e) pairs for the learned PC
  bool prepared = false;
  int matvecs = 0;
  int mode_precision = 0;
  // after the solve (reparametrization, area / volume correction)
  DBuf repw, repwn, ru, rxc, ruc, r3, rdt, rE, a0, v0, av4, qwu;
  std::vector<double> h_a0, h_v0;
  bool have_av0 = false;
  int rep_iters = 0, av_iters = 0;
  int last_iters = 0;              // iterations of the last converged solve
  // previous solution for the extrapolated initial guess (extrapolate = 1);
  // solved counts the steps since the state was last set
  DBuf xprev, xcur;
  int solved = 0;
  // inner solves of the preconditioner (inner_max_iter > 0)
  DBuf inner_work;                 // BiCGStab scratch of the inner solves, 8 n
  DBuf solve_work;                 // BiCGStab scratch of the outer solve, 8 n
  DBuf gmres_work;                 // FGMRES scratch of the outer solve, (2 m + 1) n
  DBuf x0s, Axr;                   // the initial guess (for the fallback); A x of a restart
  // viscosity contrast (VesicleProperties::update, src/VesicleProps.cc:28-52):
  // vel_coeff = (1 + lambda) / 2, dl_coeff = 1 - lambda per vesicle
  bool contrast = false;
  DBuf vcoef, dcoef, dd;
  apply_fn inner_pc = nullptr;     // the inner solve's own preconditioner
  int inner_iters = 0, local_mv = 0;

  ~Stepper() {
    if (cinfo) (void)hipFree(cinfo);
  }

  int init(const ves3d_step_params &p_) {
    prm = p_;
    const int p = prm.p;
    if (p < 2 || p > 64 || prm.nves < 1) return VES3D_INVALID_PARAMETER;
    if (prm.inner_max_iter < 0 || prm.inner_tol < 0 || prm.inner_tol >= 1 ||
        (prm.extrapolate != 0 && prm.extrapolate != 1) ||
        (prm.far_coarse != 0.0 && !(prm.far_coarse > 1.0)) ||
        (prm.far_coarse2 != 0.0 && !(prm.far_coarse > 1.0 && prm.far_coarse2 > prm.far_coarse)) ||
        (prm.fallback != 0 && prm.fallback != 1) || (prm.solver != 0 && prm.solver != 1) ||
        prm.restart < 0 || prm.restart >= kMaxVecCoefs)
      return VES3D_INVALID_PARAMETER;
    if (prm.p_up <= 0) prm.p_up = 2 * p;
    if (prm.filter_freq <= 0) prm.filter_freq = 2 * p / 3;
    nv = (long)prm.nves;
    ng = sh_ngrid(p);
    nc = sh_ncoef(p);
    n = 4 * nv * nc;
    V3D_CHECK(forces.init(p, prm.p_up, prm.filter_freq, nv, prm.bending_modulus));
    if (prm.repul_dist == 0.0) prm.repul_dist = 5e-2;  // src/Parameters.cc:46
    V3D_CHECK(stokes.init(p, prm.p_up, nv, prm.precision, prm.near_off ? 0 : 1, prm.repul_dist,
                          prm.r_near, prm.far_coarse, prm.far_coarse2));
    for (DBuf *b : {&X, &U, &f, &f2, &Sf, &ug}) V3D_CHECK(b->alloc(3 * nv * ng));
    for (DBuf *b : {&sig, &ten, &sg}) V3D_CHECK(b->alloc(nv * ng));
    V3D_CHECK(cf.alloc(3 * nv * nc));
    V3D_CHECK(rhs.alloc(n));
    V3D_CHECK(xsol.alloc(n));
    V3D_CHECK(yres.alloc(n));
    V3D_CHECK(tmpD.alloc(n));
    V3D_CHECK(frames.alloc(9 * nv));
    {
      std::vector<double> gx, gw, fg(4 * ng);
      gauss_legendre(p + 1, gx, gw);
      for (int i = 0; i <= p; ++i)
        for (int j = 0; j < 2 * p; ++j) {
          const double s = std::sqrt(1.0 - gx[i] * gx[i]), ph = M_PI * j / p;
          const long k = (long)i * 2 * p + j;
          fg[k] = s * std::cos(ph);
          fg[ng + k] = s * std::sin(ph);
          fg[2 * ng + k] = gx[i];
          fg[3 * ng + k] = gw[i];
        }
      V3D_CHECK(fgrid.alloc(4 * ng));
      V3D_HIP(hipMemcpy(fgrid.p, fg.data(), sizeof(double) * 4 * ng, hipMemcpyHostToDevice));
    }
    V3D_CHECK(post_init());
    V3D_HIP(hipMemset(U.p, 0, sizeof(double) * U.n));
    V3D_HIP(hipMemset(sig.p, 0, sizeof(double) * sig.n));
    std::vector<int2> ci(nc);
    for (int j = 0; j < 2 * p; ++j) {
      const int off = sh_blockoff(p, j), len = sh_blocklen(p, j);
      for (int k = 0; k < len; ++k) ci[off + k] = make_int2(off, len);
    }
    V3D_HIP(hipMalloc(&cinfo, sizeof(int2) * nc));
    V3D_HIP(hipMemcpy(cinfo, ci.data(), sizeof(int2) * nc, hipMemcpyHostToDevice));
    // ConfigurePrecond (src/InterfacialVelocity.cc:664-713)
    std::vector<double> bp(nc), tp(nc);
    for (int j = 0; j < 2 * p; ++j) {
      const int m = (j + 1) / 2, off = sh_blockoff(p, j), len = sh_blocklen(p, j);
      for (int k = 0; k < len; ++k) {
        const int nn = m + k;
        double b = 1.0 / std::fabs(1.0 - prm.dt * nn * nn * nn);
        b = std::fabs(b) < 1e3 ? b : 1.0;
        bp[off + k] = std::fabs(b) > 1e-10 ? b : 1.0;
        double e = 1.0;
        if (nn > 0) e = (4.0 * nn * nn - 1) * (2 * nn + 3) / (nn + 1) / (2.0 * nn * nn + 2 * nn - 1) / nn;
        tp[off + k] = std::fabs(e) > 1e-10 ? e : 1.0;
      }
    }
    V3D_CHECK(bprec.alloc(nc));
    V3D_CHECK(tprec.alloc(nc));
    V3D_HIP(hipMemcpy(bprec.p, bp.data(), sizeof(double) * nc, hipMemcpyHostToDevice));
    V3D_HIP(hipMemcpy(tprec.p, tp.data(), sizeof(double) * nc, hipMemcpyHostToDevice));
    return VES3D_SUCCESS;
  }

  int pack(const double *u, const double *s, double *x) {  // grid -> Galerkin vector
    V3D_CHECK(forces.S.forward(u, 3 * nv, cf, st));
    k_fm_to_om<<<gb(3 * nv * nc), 256, 0, st>>>(cf, 3 * nv, (int)nc, cinfo, x);
    V3D_CHECK(launch_status());
    V3D_CHECK(forces.S.forward(s, nv, cf, st));
    k_fm_to_om<<<gb(nv * nc), 256, 0, st>>>(cf, nv, (int)nc, cinfo, x + 3 * nv * nc);
    return launch_status();
  }
  int unpack(const double *x, double *u, double *s) {
    k_om_to_fm<<<gb(3 * nv * nc), 256, 0, st>>>(x, 3 * nv, (int)nc, cinfo, cf);
    V3D_CHECK(launch_status());
    V3D_CHECK(forces.S.backward(cf, 3 * nv, 0, 0, u, st));
    k_om_to_fm<<<gb(nv * nc), 256, 0, st>>>(x + 3 * nv * nc, nv, (int)nc, cinfo, cf);
    V3D_CHECK(launch_status());
    return forces.S.backward(cf, nv, 0, 0, s, st);
  }

  int prepare() {
    V3D_CHECK(forces.update(X, st));
    V3D_CHECK(stokes.setup(X, st));
    k_frames<<<(unsigned)nv, 256, 0, st>>>(X, ng, fgrid, frames);
    V3D_CHECK(launch_status());
    prepared = true;
    return VES3D_SUCCESS;
  }

  // ImplicitMatvecPhysical (src/InterfacialVelocity.cc:921-979), velocity form;
  // local: every vesicle's own operator (self-interaction only), which gives
  // the learned preconditioner its training pairs
  int matvec(const double *x, double *y, bool local = false) {
    if (!prepared) return VES3D_INVALID_PARAMETER;
    if (local && !contrast) return matvec_local_c(x, y);
    if (!local) ++matvecs;
    V3D_CHECK(unpack(x, ug, sg));
    V3D_CHECK(dev_axpby(3 * nv * ng, prm.dt, ug, 0.0, f2, st));  // Du = dt u
    // implicitTractionJump (src/InterfacialForce.cc:172-184)
    V3D_CHECK(forces.linear_bending(f2, f, st));
    V3D_CHECK(forces.gravity_add(f2, f, st));
    V3D_CHECK(forces.tensile(sg, f2, st));
    V3D_CHECK(dev_axpby(3 * nv * ng, 1.0, f2, 1.0, f, st));
    if (contrast) {  // SetDensityDL(dl_coeff u) (src/InterfacialVelocity.cc:945-951)
      k_ves_scale<<<gb(3 * nv * ng), 256, 0, st>>>(ug, dcoef, nv, 3 * ng, dd);
      V3D_CHECK(launch_status());
    }
    V3D_CHECK(stokes.apply(f, Sf, st, false, local, contrast ? dd.p : nullptr));
    V3D_CHECK(forces.S.div(Sf, ten, st));
    V3D_CHECK(dev_axpby(nv * ng, 0.0, ten, -1.0, ten, st));      // ten = -div
    if (contrast) {  // vox = vel_coeff u - Sf (:965-968)
      k_ves_axmy<<<gb(3 * nv * ng), 256, 0, st>>>(ug, vcoef, Sf, nv, 3 * ng);
      V3D_CHECK(launch_status());
    } else {
      V3D_CHECK(dev_axpby(3 * nv * ng, -1.0, Sf, 1.0, ug, st));  // vox = u - Sf
    }
    return pack(ug, ten, y);
  }

  // Every vesicle's own operator (the self interaction alone), the same map
  // as matvec(x, y, true) carried in SH coefficients: the Galerkin vector
  // holds the coefficients of u and sigma, so the forces start from them
  // (no synthesis and re-analysis of dt u and sigma), the self velocity
  // stays in coefficients, the tension row takes the divergence's low-passed
  // coefficients and the velocity row is u - S f coefficient by coefficient:
  // the same values up to rounding, five transforms fewer (an inner solve
  // applies it at every iteration).
  int matvec_local_c(const double *x, double *y) {
    if (!lcu.p)
      for (DBuf *b : {&lcu, &lcd, &lcv}) V3D_CHECK(b->alloc(3 * nv * nc));
    if (!lcs.p)
      for (DBuf *b : {&lcs, &lct}) V3D_CHECK(b->alloc(nv * nc));
    k_om_to_fm<<<gb(3 * nv * nc), 256, 0, st>>>(x, 3 * nv, (int)nc, cinfo, lcu);
    V3D_CHECK(launch_status());
    k_om_to_fm<<<gb(nv * nc), 256, 0, st>>>(x + 3 * nv * nc, nv, (int)nc, cinfo, lcs);
    V3D_CHECK(launch_status());
    V3D_CHECK(dev_axpby(3 * nv * nc, prm.dt, lcu, 0.0, lcd, st));  // Du = dt u
    // implicitTractionJump (src/InterfacialForce.cc:172-184)
    V3D_CHECK(forces.linear_bending_c(lcd, f, st));
    if (forces.has_rho) {  // the gravity traction needs dt u on the grid
      V3D_CHECK(forces.S.backward(lcd, 3 * nv, 0, 0, f2, st));
      V3D_CHECK(forces.gravity_add(f2, f, st));
    }
    V3D_CHECK(forces.tensile_c(lcs, f2, st));
    V3D_CHECK(dev_axpby(3 * nv * ng, 1.0, f2, 1.0, f, st));
    V3D_CHECK(stokes.self_coef(f, lcv, st));
    V3D_CHECK(forces.S.div_coef(lcv, ten, lct, st));                  // div (S f), low-passed
    V3D_CHECK(dev_axpby(3 * nv * nc, -1.0, lcv, 1.0, lcu, st));      // u - S f
    V3D_CHECK(dev_axpby(nv * nc, -1.0, lct, 0.0, lct, st));          // -div (S f)
    k_fm_to_om<<<gb(3 * nv * nc), 256, 0, st>>>(lcu, 3 * nv, (int)nc, cinfo, y);
    V3D_CHECK(launch_status());
    k_fm_to_om<<<gb(nv * nc), 256, 0, st>>>(lct, nv, (int)nc, cinfo, y + 3 * nv * nc);
    return launch_status();
  }

  // AssembleRhsVel (src/InterfacialVelocity.cc:715-787): the explicit
  // traction jump plus the repulsion force (:733)
  int assemble_rhs(double *b) {
    if (!prepared) return VES3D_INVALID_PARAMETER;
    k_shear<<<gb(nv * ng), 256, 0, st>>>(X, nv, ng, prm.shear_rate, ug);
    V3D_CHECK(launch_status());
    V3D_CHECK(forces.bending(f, st));  // explicitTractionJump (src/InterfacialForce.cc:162-170)
    V3D_CHECK(forces.gravity_add(X, f, st));
    V3D_CHECK(stokes.apply(f, Sf, st, true));
    V3D_CHECK(dev_axpby(3 * nv * ng, 1.0, Sf, 1.0, ug, st));
    V3D_CHECK(forces.S.div(ug, ten, st));
    return pack(ug, ten, b);
  }

  // AssembleRhsVel into b and the matvec y = A x together: the two Stokes
  // evaluations share one far-field pass (two densities, same sources)
  int rhs_and_matvec(double *b, const double *x, double *y) {
    if (!prepared) return VES3D_INVALID_PARAMETER;
    if (contrast) {  // the matvec carries a double layer the right-hand side has not
      V3D_CHECK(assemble_rhs(b));
      return matvec(x, y);
    }
    if (!ub.p)
      for (DBuf *d : {&ub, &fb, &Sfb}) V3D_CHECK(d->alloc(3 * nv * ng));
    ++matvecs;
    // density of A x (as in matvec)
    V3D_CHECK(unpack(x, ug, sg));
    V3D_CHECK(dev_axpby(3 * nv * ng, prm.dt, ug, 0.0, f2, st));
    V3D_CHECK(forces.linear_bending(f2, f, st));
    V3D_CHECK(forces.gravity_add(f2, f, st));
    V3D_CHECK(forces.tensile(sg, f2, st));
    V3D_CHECK(dev_axpby(3 * nv * ng, 1.0, f2, 1.0, f, st));
    // density of the right-hand side (as in assemble_rhs)
    k_shear<<<gb(nv * ng), 256, 0, st>>>(X, nv, ng, prm.shear_rate, ub);
    V3D_CHECK(launch_status());
    V3D_CHECK(forces.bending(fb, st));
    V3D_CHECK(forces.gravity_add(X, fb, st));
    V3D_CHECK(stokes.apply2(fb, Sfb, true, f, Sf, st));
    // right-hand side
    V3D_CHECK(dev_axpby(3 * nv * ng, 1.0, Sfb, 1.0, ub, st));
    V3D_CHECK(forces.S.div(ub, ten, st));
    V3D_CHECK(pack(ub, ten, b));
    // matvec
    V3D_CHECK(forces.S.div(Sf, ten, st));
    V3D_CHECK(dev_axpby(nv * ng, 0.0, ten, -1.0, ten, st));      // ten = -div
    V3D_CHECK(dev_axpby(3 * nv * ng, -1.0, Sf, 1.0, ug, st));    // vox = u - Sf
    return pack(ug, ten, y);
  }

  // ImplicitPrecond (src/InterfacialVelocity.cc:1054-1109), Galerkin form
  int precond_diag(const double *x, double *y) {
    k_scale_om<<<gb(3 * nv * nc), 256, 0, st>>>(x, 3 * nv, (int)nc, cinfo, bprec, y);
    V3D_CHECK(launch_status());
    k_scale_om<<<gb(nv * nc), 256, 0, st>>>(x + 3 * nv * nc, nv, (int)nc, cinfo, tprec,
                                            y + 3 * nv * nc);
    return launch_status();
  }

  int true_residual(const double *x, const double *b, double *rel) {
    V3D_CHECK(matvec(x, yres));
    V3D_CHECK(dev_axpby(n, 1.0, b, -1.0, yres, st));
    int e1 = 0, e2 = 0;
    const double nr = dev_norm(n, yres, st, &e1), nb = dev_norm(n, b, st, &e2);
    V3D_CHECK(e1);
    V3D_CHECK(e2);
    *rel = nr / (nb == 0 ? 1.0 : nb);
    return VES3D_SUCCESS;
  }

  static int A_cb(const void *ctx, const double *x, double *y) {
    Stepper *s = (Stepper *)ctx;
    V3D_CHECK(s->matvec(x, y));
    if (s->collect && s->lpc) {
      // training pair from the vesicles' own operators: input D A_loc x,
      // target x (the far-field coupling would enter a per-vesicle
      // regression as noise correlated with its input)
      if (!s->yloc.p) V3D_CHECK(s->yloc.alloc(s->n));
      V3D_CHECK(s->matvec(x, s->yloc, true));
      V3D_CHECK(s->precond_diag(s->yloc, s->tmpD));
      V3D_CHECK(s->lpc->record(s->tmpD, x, s->nv, s->prm.p, s->cinfo, s->frames, s->st));
    }
    return VES3D_SUCCESS;
  }
  static int Pdiag_cb(const void *ctx, const double *x, double *y) {
    return ((Stepper *)ctx)->precond_diag(x, y);
  }
  // every vesicle's own operator (no far field, no near correction)
  static int Aloc_cb(const void *ctx, const double *x, double *y) {
    Stepper *s = (Stepper *)ctx;
    ++s->local_mv;
    return s->matvec(x, y, true);
  }
  // Preconditioner with an inner solve: y ~ A_loc^-1 x, the block-diagonal
  // (per-vesicle) operator inverted by BiCGStab from y = 0 to inner_tol,
  // preconditioned by inner_pc (the diagonal or the learned preconditioner).
  // The result is A_loc^-1 x to relative accuracy inner_tol, so the outer
  // BiCGStab sees a preconditioner that departs from a fixed linear map by
  // that much; the true-residual check after the solve judges the solution.
  static int Pinner_cb(const void *ctx, const double *x, double *y) {
    Stepper *s = (Stepper *)ctx;
    if (!s->inner_work.p) V3D_CHECK(s->inner_work.alloc(8 * s->n));
    V3D_CHECK(dev_fill(s->n, 0.0, y, s->st));
    int it = s->prm.inner_max_iter;
    double tol = s->prm.inner_tol > 0 ? s->prm.inner_tol : 1e-3;
    int status = 0;
    V3D_CHECK(bicgstab(s->n, Aloc_cb, s, s->inner_pc, s, y, x, &it, &tol, &status, s->st,
                       nullptr, s->inner_work, true));
    s->inner_iters += it;
    return VES3D_SUCCESS;
  }
  static int Plearn_cb(const void *ctx, const double *x, double *y) {
    Stepper *s = (Stepper *)ctx;
    V3D_CHECK(s->precond_diag(x, y));
    return s->lpc->apply_inplace(y, s->nv, s->prm.p, s->prm.filter_freq, s->cinfo, s->frames,
                                 s->st);
  }

  // ---- reparametrization and area / volume correction ---------------------
  // Defaults of the reference's Parameters (src/Parameters.cc:39-45):
  // PolyKReparam with exponent 4, rep_ts = 1e-3, rep_tol = 1e-4 (FP64),
  // rep_maxit = 10, no upsampling; the area / volume tolerance 1e-12
  // (include/EvolveSurface.h:105).
  static constexpr int kRepExp = 4, kRepMaxit = 10;
  static constexpr double kRepTs = 1e-3, kRepTol = 1e-4, kAvTol = 1e-12;

  int post_init() {
    const int p = prm.p, pu = prm.p_up;
    const long mu = sh_ngrid(pu);
    // A[n] (src/InterfacialVelocity.cc:1553-1566) per coefficient
    std::vector<double> w(nc), wn(nc);
    const long ff = kRepExp ? p / 2 : p / 3;
    for (int j = 0; j < 2 * p; ++j) {
      const int m = (j + 1) / 2, off = sh_blockoff(p, j), len = sh_blocklen(p, j);
      for (int k = 0; k < len; ++k) {
        const int deg = m + k;
        double a = 1.0 - (kRepExp ? std::pow(deg * 1.0 / ff, kRepExp) : 0.0);
        a *= deg > ff ? 0.0 : 1.0;
        w[off + k] = 1.0 - a;
        wn[off + k] = -(1.0 - a);
      }
    }
    std::vector<double> q;
    legendre_weights(pu, q);
    V3D_CHECK(repw.alloc(nc));
    V3D_CHECK(repwn.alloc(nc));
    V3D_CHECK(qwu.alloc(pu + 1));
    V3D_HIP(hipMemcpy(repw.p, w.data(), sizeof(double) * nc, hipMemcpyHostToDevice));
    V3D_HIP(hipMemcpy(repwn.p, wn.data(), sizeof(double) * nc, hipMemcpyHostToDevice));
    V3D_HIP(hipMemcpy(qwu.p, q.data(), sizeof(double) * (pu + 1), hipMemcpyHostToDevice));
    V3D_CHECK(ru.alloc(3 * nv * ng));
    for (DBuf *b : {&rxc, &ruc}) V3D_CHECK(b->alloc(3 * nv * nc));
    V3D_CHECK(r3.alloc(3 * nv));
    V3D_CHECK(rdt.alloc(nv));
    V3D_CHECK(rE.alloc(2 * nv));
    V3D_CHECK(a0.alloc(nv));
    V3D_CHECK(v0.alloc(nv));
    V3D_CHECK(av4.alloc(4 * nv));
    (void)mu;
    return VES3D_SUCCESS;
  }

  // area / volume of every vesicle on the upsampled surface into av4 (and host)
  int av_measure(std::vector<double> &h4) {
    const long mu = sh_ngrid(prm.p_up);
    V3D_CHECK(forces.resample(X, 3 * nv, prm.p, prm.p_up, forces.xu, st));
    V3D_CHECK(forces.Su.update(forces.xu, st));
    k_av_integrals<<<(unsigned)nv, 256, 0, st>>>(forces.xu, forces.Su.nrm, forces.Su.W,
                                                 forces.Su.H, qwu, mu, 2 * prm.p_up, av4);
    V3D_CHECK(launch_status());
    h4.resize(4 * nv);
    V3D_HIP(hipMemcpyAsync(h4.data(), av4.p, sizeof(double) * 4 * nv, hipMemcpyDeviceToHost, st));
    return hip_status(hipStreamSynchronize(st));
  }

  // the reference area / volume (EvolveSurface::Evolve, src/EvolveSurface.cc:119-126)
  int record_area_volume() {
    std::vector<double> h4;
    V3D_CHECK(av_measure(h4));
    h_a0.resize(nv);
    h_v0.resize(nv);
    for (long v = 0; v < nv; ++v) {
      h_a0[v] = h4[4 * v];
      h_v0[v] = h4[4 * v + 1];
    }
    V3D_HIP(hipMemcpy(a0.p, h_a0.data(), sizeof(double) * nv, hipMemcpyHostToDevice));
    V3D_HIP(hipMemcpy(v0.p, h_v0.data(), sizeof(double) * nv, hipMemcpyHostToDevice));
    have_av0 = true;
    return VES3D_SUCCESS;
  }

  // InterfacialVelocity::reparam, no upsampling (src/InterfacialVelocity.cc:1639-1773)
  int reparam() {
    SurfaceG &S = forces.S;
    int it = 0;
    int have_E = 0;
    std::vector<double> hdt(nv);
    while (it < kRepMaxit) {
      V3D_CHECK(S.update(X, st));
      V3D_CHECK(S.forward(X, 3 * nv, rxc, st));
      V3D_HIP(hipMemcpyAsync(ruc.p, rxc.p, sizeof(double) * 3 * nv * nc, hipMemcpyDeviceToDevice, st));
      k_scale_fm<<<gb(3 * nv * nc), 256, 0, st>>>(ruc, 3 * nv, (int)nc, repwn);
      V3D_CHECK(launch_status());
      V3D_CHECK(S.backward(ruc, 3 * nv, 0, 0, ru, st));            // sht_filter
      for (int r = 0; r < 2; ++r) {                                 // tangent + band-limit, twice
        k_tangent<<<gb(nv * ng), 256, 0, st>>>(ru, S.nrm, nv, ng);
        V3D_CHECK(launch_status());
        V3D_CHECK(S.forward(ru, 3 * nv, ruc, st));
        V3D_CHECK(S.backward(ruc, 3 * nv, 0, 0, ru, st));
      }
      k_rep_norm_inner<<<(unsigned)nv, 256, 0, st>>>(ru, ng, rxc, ruc, (int)nc, repw, r3);
      V3D_CHECK(launch_status());
      k_rep_step<<<gb(nv), 256, 0, st>>>(r3, nv, kRepTs, kRepTol, have_E, rE, rdt);
      V3D_CHECK(launch_status());
      V3D_HIP(hipMemcpyAsync(hdt.data(), rdt.p, sizeof(double) * nv, hipMemcpyDeviceToHost, st));
      V3D_HIP(hipStreamSynchronize(st));
      double dt_max = 0;
      for (double d : hdt) dt_max = std::max(dt_max, d);
      if (dt_max == 0) break;
      V3D_HIP(hipMemcpyAsync(rE.p, rE.p + nv, sizeof(double) * nv, hipMemcpyDeviceToDevice, st));
      have_E = 1;
      k_rep_update<<<gb(3 * nv * ng), 256, 0, st>>>(X, ru, rdt, nv, 3 * ng);
      V3D_CHECK(launch_status());
      ++it;
    }
    rep_iters = it;
    return hip_status(hipStreamSynchronize(st));
  }

  // EvolveSurface::AreaVolumeCorrection (src/EvolveSurface.cc:352-470)
  int area_volume_correction() {
    if (!have_av0) V3D_CHECK(record_area_volume());
    const long mu = sh_ngrid(prm.p_up);
    std::vector<double> h4;
    int it = -1;
    while (++it < kRepMaxit) {
      V3D_CHECK(av_measure(h4));
      double amax = 0, vmax = 0, aerr = 0, verr = 0;
      for (long v = 0; v < nv; ++v) {
        amax = std::max(amax, std::fabs(h_a0[v]));
        vmax = std::max(vmax, std::fabs(h_v0[v]));
        aerr = std::max(aerr, std::fabs(h_a0[v] - h4[4 * v]));
        verr = std::max(verr, std::fabs(h_v0[v] - h4[4 * v + 1]));
      }
      if (std::max(aerr / amax, verr / vmax) < kAvTol) break;
      k_av_update<<<gb(nv * mu), 256, 0, st>>>(forces.xu, forces.Su.nrm, forces.Su.H, av4, a0, v0,
                                               nv, mu);
      V3D_CHECK(launch_status());
      V3D_CHECK(forces.resample(forces.xu, 3 * nv, prm.p_up, prm.p, X, st));
    }
    av_iters = it;
    return hip_status(hipStreamSynchronize(st));
  }

  // A solution is accepted when BiCGStab converged and its true residual
  // is within 5% of the tolerance.
  static constexpr double kTrueResSlack = 1.05;
  static constexpr int kDefaultRestart = 30;  // FGMRES restart length
  static constexpr int kRestarts = 2;

  // BiCGStab on the step's system from xsol (whose A x is Ax0), then the
  // true residual of the result (src/InterfacialVelocity.cc:181-188).  When
  // BiCGStab reports convergence but the true residual misses the
  // tolerance (the recursive residual drifted from it), BiCGStab continues
  // from the current solution, at most kRestarts times, within the same
  // iteration budget.  it: iterations used in total.
  int solve(apply_fn P, int budget, int *it, double *tol, int *status, double *true_rel) {
    if (prm.solver == 1) {  // FGMRES: its residual is formed explicitly, no check matvec
      const int m = prm.restart > 0 ? prm.restart : kDefaultRestart;
      if (!gmres_work.p) V3D_CHECK(gmres_work.alloc((2L * m + 1) * n));
      int k = budget;
      double t = prm.tol;
      V3D_CHECK(fgmres(n, A_cb, this, P, this, xsol, rhs, &k, &t, status, st, m, Ax0,
                       gmres_work));
      *it = k;
      *tol = t;
      *true_rel = t;
      return VES3D_SUCCESS;
    }
    const double *ax = Ax0;
    int used = 0;
    for (int r = 0;; ++r) {
      int k = budget - used;
      double t = prm.tol;
      V3D_CHECK(bicgstab(n, A_cb, this, P, this, xsol, rhs, &k, &t, status, st, ax, solve_work));
      used += k;
      *tol = t;
      V3D_CHECK(true_residual(xsol, rhs, true_rel));  // yres = b - A xsol
      if (*status != 100 || *true_rel <= prm.tol * kTrueResSlack || r == kRestarts ||
          used >= budget)
        break;
      if (!Axr.p) V3D_CHECK(Axr.alloc(n));
      V3D_CHECK(dev_axpy3(n, -1.0, yres, rhs, Axr, st));  // A xsol = b - (b - A xsol)
      ax = Axr;
    }
    *it = used;
    return VES3D_SUCCESS;
  }

  int step(ves3d_step_stats *stats) {
    using clk = std::chrono::steady_clock;
    auto t0 = clk::now();
    stokes.timing_begin();
    matvecs = 0;
    V3D_CHECK(prepare());
    V3D_HIP(hipStreamSynchronize(st));
    auto t1 = clk::now();
    V3D_CHECK(pack(U, sig, xsol));  // AssembleInitial (src/InterfacialVelocity.cc:873-919)
    if (prm.extrapolate && solved >= 1) {
      // (U, sigma) is the last step's solution x_{n-1}: start from
      // 2 x_{n-1} - x_{n-2} once two solutions exist, and keep x_{n-1}
      if (!xprev.p) V3D_CHECK(xprev.alloc(n));
      if (!xcur.p) V3D_CHECK(xcur.alloc(n));
      V3D_CHECK(dev_copy(n, xsol, xcur, st));
      if (solved >= 2) V3D_CHECK(dev_axpby(n, -1.0, xprev, 2.0, xsol, st));
      V3D_CHECK(dev_copy(n, xcur, xprev, st));
    }
    // the right-hand side and the solver's first matvec A x0 in one pass
    if (!Ax0.p) V3D_CHECK(Ax0.alloc(n));
    V3D_CHECK(rhs_and_matvec(rhs, xsol, Ax0));
    V3D_HIP(hipStreamSynchronize(st));
    auto t2 = clk::now();
    apply_fn P = nullptr;
    const bool learned = prm.precond == 2 && lpc && lpc->ready();
    if (prm.precond == 1 || (prm.precond == 2 && !learned)) P = Pdiag_cb;
    if (learned) P = Plearn_cb;
    inner_iters = local_mv = 0;
    if (prm.inner_max_iter > 0 && P) {
      inner_pc = P;
      P = Pinner_cb;
    }
    if (!solve_work.p) V3D_CHECK(solve_work.alloc(8 * n));
    if (learned) {  // the fallback restarts from the same initial guess, whose A x is Ax0
      if (!x0s.p) V3D_CHECK(x0s.alloc(n));
      V3D_CHECK(dev_copy(n, xsol, x0s, st));
    }
    // a learned solve gets an iteration budget tied to the last solve's count
    // (a third of max_iter before there is one), so that a preconditioner
    // that does not deliver costs little before the fallback below
    const int budget = !learned ? prm.max_iter
                       : last_iters > 0 ? std::min(prm.max_iter, 2 * last_iters + 10)
                                        : std::max(10, prm.max_iter / 3);
    int it = 0, status = 0;
    double tol = prm.tol, true_rel = 0;
    V3D_CHECK(solve(P, budget, &it, &tol, &status, &true_rel));
    bool fell_back = false;
    if (learned && !(status == 100 && true_rel <= prm.tol * kTrueResSlack)) {
      // the learned preconditioner did not deliver: BiCGStab starts again
      // from the initial guess with the diagonal spectral preconditioner
      // (ImplicitPrecond), or unpreconditioned
      fell_back = true;
      V3D_CHECK(dev_copy(n, x0s, xsol, st));
      tol = prm.tol;
      V3D_CHECK(solve(prm.fallback == 1 ? nullptr : Pdiag_cb, prm.max_iter, &it, &tol, &status,
                      &true_rel));
    }
    // src/InterfacialVelocity.cc:176-188: a solve that failed, or whose
    // true residual misses the tolerance, is a DivergenceError; the state
    // (positions, velocity, tension) is then left as it was
    const bool converged = status == 100 && true_rel <= prm.tol * kTrueResSlack;
    if (converged) {
      last_iters = it;
      ++solved;
    } else {
      solved = 0;  // no solution history to extrapolate from
    }
    auto t3 = clk::now();
    if (converged) {
      V3D_CHECK(unpack(xsol, U, sig));
      V3D_CHECK(dev_axpby(3 * nv * ng, prm.dt, U, 1.0, X, st));  // dx = dt u
    }
    V3D_HIP(hipStreamSynchronize(st));
    auto t3b = clk::now();
    rep_iters = av_iters = 0;
    if (!prm.post_off && converged) {  // EvolveSurface::Evolve after the update (src/EvolveSurface.cc:330-335)
      V3D_CHECK(reparam());
      V3D_CHECK(area_volume_correction());
    }
    auto t4 = clk::now();
    double far_ms = 0, far2_ms = 0;
    int far_calls = 0, far2_calls = 0;
    double near_ms = 0;
    V3D_CHECK(stokes.far_time(&far_ms, &far_calls, &far2_ms, &far2_calls, &near_ms));
    prepared = false;
    if (stats) {
      auto ms = [](clk::time_point a, clk::time_point b) {
        return std::chrono::duration<double, std::milli>(b - a).count();
      };
      stats->iterations = it;
      stats->matvecs = matvecs;
      stats->status = status;
      stats->relres = tol;
      stats->true_relres = true_rel;
      stats->fell_back = fell_back ? 1 : 0;
      stats->ms_setup = ms(t0, t1);
      stats->ms_rhs = ms(t1, t2);
      stats->ms_solve = ms(t2, t3);
      stats->ms_total = ms(t0, t4);
      stats->ms_far = far_ms;
      stats->far_calls = far_calls;
      stats->near_pairs = stokes.use_near ? (int)stokes.near.npairs : 0;
      stats->r_near = stokes.use_near ? stokes.near.r_near : 0.0;
      stats->reparam_iters = rep_iters;
      stats->av_iters = av_iters;
      stats->ms_post = ms(t3b, t4);
      stats->ms_far2 = far2_ms;
      stats->far2_calls = far2_calls;
      stats->inner_iters = inner_iters;
      stats->local_matvecs = local_mv;
      stats->far_pairs = stokes.far_pairs;
      stats->ms_near = near_ms;
    }
    return converged ? VES3D_SUCCESS : VES3D_DIVERGENCE_ERROR;
  }
};

}  // namespace v3d

// ---------------------------------------------------------------- C ABI
using namespace v3d;

extern "C" {

int ves3d_hip_bicgstab(size_t n, ves3d_apply_fn A, const void *A_ctx, ves3d_apply_fn P,
                       const void *P_ctx, double *x, const double *b, int *max_iter,
                       double *tol, int *bicgs_status, ves3d_stream_t stream) {
  if (!A || !x || !b || !max_iter || !tol || !bicgs_status) return VES3D_INVALID_PARAMETER;
  return bicgstab((long)n, (apply_fn)A, A_ctx, (apply_fn)P, P_ctx, x, b, max_iter, tol,
                  bicgs_status, (hipStream_t)stream);
}

int ves3d_hip_fgmres(size_t n, ves3d_apply_fn A, const void *A_ctx, ves3d_apply_fn P,
                     const void *P_ctx, double *x, const double *b, int *max_iter, double *tol,
                     int *status, int restart, ves3d_stream_t stream) {
  if (!A || !x || !b || !max_iter || !tol || !status || n < 1 || restart < 1 ||
      restart >= kMaxVecCoefs)
    return VES3D_INVALID_PARAMETER;
  return fgmres((long)n, (apply_fn)A, A_ctx, (apply_fn)P, P_ctx, x, b, max_iter, tol, status,
                (hipStream_t)stream, restart);
}

int ves3d_hip_stepper_create(const ves3d_step_params *prm, void **handle) {
  if (!prm || !handle) return VES3D_INVALID_PARAMETER;
  Stepper *s = new Stepper();
  int e = s->init(*prm);
  if (e != VES3D_SUCCESS) {
    delete s;
    return e;
  }
  *handle = s;
  return VES3D_SUCCESS;
}

int ves3d_hip_stepper_destroy(void *h) {
  delete (Stepper *)h;
  return VES3D_SUCCESS;
}

int ves3d_hip_stepper_size(void *h, size_t *n) {
  if (!h || !n) return VES3D_INVALID_PARAMETER;
  *n = (size_t)((Stepper *)h)->n;
  return VES3D_SUCCESS;
}

int ves3d_hip_stepper_set_state(void *h, const double *X, const double *U, const double *sigma) {
  if (!h) return VES3D_INVALID_PARAMETER;
  Stepper *s = (Stepper *)h;
  if (X) V3D_HIP(hipMemcpy(s->X.p, X, sizeof(double) * s->X.n, hipMemcpyDeviceToDevice));
  if (U) V3D_HIP(hipMemcpy(s->U.p, U, sizeof(double) * s->U.n, hipMemcpyDeviceToDevice));
  if (sigma) V3D_HIP(hipMemcpy(s->sig.p, sigma, sizeof(double) * s->sig.n, hipMemcpyDeviceToDevice));
  s->prepared = false;
  s->solved = 0;  // a new trajectory: no solution history to extrapolate
  if (X) V3D_CHECK(s->record_area_volume());  // the targets of the area / volume correction
  return VES3D_SUCCESS;
}

int ves3d_hip_stepper_get_state(void *h, double *X, double *U, double *sigma) {
  if (!h) return VES3D_INVALID_PARAMETER;
  Stepper *s = (Stepper *)h;
  if (X) V3D_HIP(hipMemcpy(X, s->X.p, sizeof(double) * s->X.n, hipMemcpyDeviceToDevice));
  if (U) V3D_HIP(hipMemcpy(U, s->U.p, sizeof(double) * s->U.n, hipMemcpyDeviceToDevice));
  if (sigma) V3D_HIP(hipMemcpy(sigma, s->sig.p, sizeof(double) * s->sig.n, hipMemcpyDeviceToDevice));
  return VES3D_SUCCESS;
}

int ves3d_hip_stepper_prepare(void *h) {
  if (!h) return VES3D_INVALID_PARAMETER;
  Stepper *s = (Stepper *)h;
  V3D_CHECK(s->prepare());
  return hip_status(hipStreamSynchronize(s->st));
}

int ves3d_hip_stepper_rhs(void *h, double *b) {
  if (!h || !b) return VES3D_INVALID_PARAMETER;
  Stepper *s = (Stepper *)h;
  V3D_CHECK(s->assemble_rhs(b));
  return hip_status(hipStreamSynchronize(s->st));
}

int ves3d_hip_stepper_matvec(void *h, const double *x, double *y) {
  if (!h || !x || !y) return VES3D_INVALID_PARAMETER;
  Stepper *s = (Stepper *)h;
  V3D_CHECK(s->matvec(x, y));
  return hip_status(hipStreamSynchronize(s->st));
}

int ves3d_hip_stepper_rhs_matvec(void *h, const double *x, double *b, double *y) {
  if (!h || !x || !b || !y) return VES3D_INVALID_PARAMETER;
  Stepper *s = (Stepper *)h;
  V3D_CHECK(s->rhs_and_matvec(b, x, y));
  return hip_status(hipStreamSynchronize(s->st));
}

int ves3d_hip_stepper_residual(void *h, double *relres_host) {
  if (!h || !relres_host) return VES3D_INVALID_PARAMETER;
  Stepper *s = (Stepper *)h;
  if (!s->prepared) V3D_CHECK(s->prepare());
  V3D_CHECK(s->assemble_rhs(s->rhs));
  V3D_CHECK(s->pack(s->U, s->sig, s->xsol));
  V3D_CHECK(s->true_residual(s->xsol, s->rhs, relres_host));
  return hip_status(hipStreamSynchronize(s->st));
}

int ves3d_hip_stepper_far_pairs(void *h, double *pairs_host) {
  if (!h || !pairs_host) return VES3D_INVALID_PARAMETER;
  Stepper *s = (Stepper *)h;
  if (!s->prepared) return VES3D_INVALID_PARAMETER;
  *pairs_host = s->stokes.far_pairs;
  return VES3D_SUCCESS;
}

int ves3d_hip_stepper_matvec_local(void *h, const double *x, double *y) {
  if (!h || !x || !y) return VES3D_INVALID_PARAMETER;
  Stepper *s = (Stepper *)h;
  V3D_CHECK(s->matvec(x, y, true));
  return hip_status(hipStreamSynchronize(s->st));
}

int ves3d_hip_stepper_precond_learned(void *h, const double *x, double *y) {
  if (!h || !x || !y) return VES3D_INVALID_PARAMETER;
  Stepper *s = (Stepper *)h;
  if (!s->prepared || !s->lpc || !s->lpc->ready()) return VES3D_INVALID_PARAMETER;
  V3D_CHECK(Stepper::Plearn_cb(s, x, y));
  return hip_status(hipStreamSynchronize(s->st));
}

int ves3d_hip_stepper_frames(void *h, double *frames) {
  if (!h || !frames) return VES3D_INVALID_PARAMETER;
  Stepper *s = (Stepper *)h;
  if (!s->prepared) return VES3D_INVALID_PARAMETER;
  V3D_HIP(hipMemcpyAsync(frames, s->frames.p, sizeof(double) * 9 * s->nv, hipMemcpyDeviceToDevice,
                         s->st));
  return hip_status(hipStreamSynchronize(s->st));
}

int ves3d_hip_stepper_precond(void *h, const double *x, double *y) {
  if (!h || !x || !y) return VES3D_INVALID_PARAMETER;
  Stepper *s = (Stepper *)h;
  V3D_CHECK(s->precond_diag(x, y));
  return hip_status(hipStreamSynchronize(s->st));
}

int ves3d_hip_stepper_precond_inner(void *h, const double *x, double *y, int learned,
                                    int *iters_host) {
  if (!h || !x || !y) return VES3D_INVALID_PARAMETER;
  Stepper *s = (Stepper *)h;
  if (!s->prepared) return VES3D_INVALID_PARAMETER;
  if (learned && !(s->lpc && s->lpc->ready())) return VES3D_INVALID_PARAMETER;
  const ves3d_step_params keep = s->prm;
  if (s->prm.inner_max_iter <= 0) s->prm.inner_max_iter = 50;
  s->inner_pc = learned ? Stepper::Plearn_cb : Stepper::Pdiag_cb;
  s->inner_iters = s->local_mv = 0;
  const int e = Stepper::Pinner_cb(s, x, y);
  s->prm = keep;
  V3D_CHECK(e);
  if (iters_host) *iters_host = s->inner_iters;
  return hip_status(hipStreamSynchronize(s->st));
}

int ves3d_hip_stepper_step(void *h, ves3d_step_stats *stats) {
  if (!h) return VES3D_INVALID_PARAMETER;
  return ((Stepper *)h)->step(stats);
}

int ves3d_hip_stepper_reparam(void *h, int *iters_host) {
  if (!h) return VES3D_INVALID_PARAMETER;
  Stepper *s = (Stepper *)h;
  V3D_CHECK(s->reparam());
  s->prepared = false;
  if (iters_host) *iters_host = s->rep_iters;
  return VES3D_SUCCESS;
}

int ves3d_hip_stepper_area_volume(void *h, int *iters_host) {
  if (!h) return VES3D_INVALID_PARAMETER;
  Stepper *s = (Stepper *)h;
  V3D_CHECK(s->area_volume_correction());
  s->prepared = false;
  if (iters_host) *iters_host = s->av_iters;
  return VES3D_SUCCESS;
}

int ves3d_hip_stepper_set_area_volume(void *h, const double *area_host, const double *vol_host) {
  if (!h || !area_host || !vol_host) return VES3D_INVALID_PARAMETER;
  Stepper *s = (Stepper *)h;
  s->h_a0.assign(area_host, area_host + s->nv);
  s->h_v0.assign(vol_host, vol_host + s->nv);
  V3D_HIP(hipMemcpy(s->a0.p, area_host, sizeof(double) * s->nv, hipMemcpyHostToDevice));
  V3D_HIP(hipMemcpy(s->v0.p, vol_host, sizeof(double) * s->nv, hipMemcpyHostToDevice));
  s->have_av0 = true;
  return VES3D_SUCCESS;
}

int ves3d_hip_stepper_set_vesicle_props(void *h, const double *bending_modulus_host,
                                        const double *viscosity_contrast_host,
                                        const double *excess_density_host,
                                        const double *gravity_field_host) {
  if (!h) return VES3D_INVALID_PARAMETER;
  Stepper *s = (Stepper *)h;
  if (viscosity_contrast_host) {
    bool con = false;
    std::vector<double> vc(s->nv), dc(s->nv);
    for (long v = 0; v < s->nv; ++v) {
      const double lam = viscosity_contrast_host[v];
      if (!(lam >= 0.0)) return VES3D_INVALID_PARAMETER;
      con |= std::fabs(lam - 1.0) > 1e-12;  // has_contrast (src/VesicleProps.cc:47-48)
      vc[v] = 0.5 + 0.5 * lam;
      dc[v] = 1.0 - lam;
    }
    if (con) {
      V3D_CHECK(s->stokes.enable_dl());
      if (!s->vcoef.p) {
        V3D_CHECK(s->vcoef.alloc(s->nv));
        V3D_CHECK(s->dcoef.alloc(s->nv));
        V3D_CHECK(s->dd.alloc(3 * s->nv * s->ng));
      }
      V3D_HIP(hipMemcpy(s->vcoef.p, vc.data(), sizeof(double) * s->nv, hipMemcpyHostToDevice));
      V3D_HIP(hipMemcpy(s->dcoef.p, dc.data(), sizeof(double) * s->nv, hipMemcpyHostToDevice));
    }
    s->contrast = con;
  }
  V3D_CHECK(s->forces.set_props(bending_modulus_host, excess_density_host, gravity_field_host));
  s->prepared = false;
  return VES3D_SUCCESS;
}

int ves3d_hip_stepper_attach_pc(void *h, void *pc, int collect) {
  if (!h) return VES3D_INVALID_PARAMETER;
  Stepper *s = (Stepper *)h;
  LearnedPC *l = (LearnedPC *)pc;
  if (l && l->num_params() <= 0) return VES3D_INVALID_PARAMETER;
  s->lpc = l;
  s->collect = (collect != 0) && l;
  return VES3D_SUCCESS;
}

int ves3d_hip_lpc_create(int nin, int width, size_t capacity, unsigned long long seed,
                         void **handle) {
  if (!handle || nin < 1 || width < 1) return VES3D_INVALID_PARAMETER;
  LearnedPC *l = new LearnedPC();
  int e = l->init(nin, width, (long)capacity, seed);
  if (e != VES3D_SUCCESS) {
    delete l;
    return e;
  }
  *handle = l;
  return VES3D_SUCCESS;
}

int ves3d_hip_lpc_destroy(void *h) {
  delete (LearnedPC *)h;
  return VES3D_SUCCESS;
}

int ves3d_hip_lpc_num_params(void *h, size_t *n) {
  if (!h || !n) return VES3D_INVALID_PARAMETER;
  *n = (size_t)((LearnedPC *)h)->num_params();
  return VES3D_SUCCESS;
}

int ves3d_hip_lpc_samples(void *h, size_t *n) {
  if (!h || !n) return VES3D_INVALID_PARAMETER;
  *n = (size_t)((LearnedPC *)h)->samples();
  return VES3D_SUCCESS;
}

int ves3d_hip_lpc_set_params(void *h, const float *params) {
  if (!h || !params) return VES3D_INVALID_PARAMETER;
  LearnedPC *l = (LearnedPC *)h;
  V3D_HIP(hipMemcpy(l->params(), params, sizeof(float) * l->num_params(), hipMemcpyDeviceToDevice));
  V3D_CHECK(l->reset_optimizer());
  l->params_changed();
  l->set_ready(true);
  return VES3D_SUCCESS;
}

int ves3d_hip_lpc_set_activations(void *h, int bf16) {
  if (!h) return VES3D_INVALID_PARAMETER;
  return ((LearnedPC *)h)->set_activations(bf16 ? 1 : 0);
}

int ves3d_hip_lpc_get_params(void *h, float *params) {
  if (!h || !params) return VES3D_INVALID_PARAMETER;
  LearnedPC *l = (LearnedPC *)h;
  return hip_status(
      hipMemcpy(params, l->params(), sizeof(float) * l->num_params(), hipMemcpyDeviceToDevice));
}

int ves3d_hip_lpc_forward(void *h, const float *Z, size_t B, float *Y, ves3d_stream_t stream) {
  if (!h || !Z || !Y || B == 0) return VES3D_INVALID_PARAMETER;
  return ((LearnedPC *)h)->forward(Z, (long)B, Y, (hipStream_t)stream);
}

int ves3d_hip_lpc_loss_grad(void *h, const float *Z, const float *T, size_t B, double *loss,
                            float *grads, ves3d_stream_t stream) {
  if (!h || !Z || !T || B == 0) return VES3D_INVALID_PARAMETER;
  LearnedPC *l = (LearnedPC *)h;
  V3D_CHECK(l->loss_grad(Z, T, (long)B, loss, (hipStream_t)stream));
  if (grads)
    V3D_HIP(hipMemcpyAsync(grads, l->grads(), sizeof(float) * l->num_params(),
                           hipMemcpyDeviceToDevice, (hipStream_t)stream));
  return hip_status(hipStreamSynchronize((hipStream_t)stream));
}

int ves3d_hip_lpc_set_weight_decay(void *h, float wd) {
  if (!h || !(wd >= 0.0f)) return VES3D_INVALID_PARAMETER;
  ((LearnedPC *)h)->set_weight_decay(wd);
  return VES3D_SUCCESS;
}

int ves3d_hip_lpc_adam(void *h, float lr, ves3d_stream_t stream) {
  if (!h) return VES3D_INVALID_PARAMETER;
  return ((LearnedPC *)h)->adam(lr, (hipStream_t)stream);
}

int ves3d_hip_lpc_train(void *h, int epochs, int batch, float lr, double *last_loss) {
  if (!h) return VES3D_INVALID_PARAMETER;
  return ((LearnedPC *)h)->train(epochs, batch, lr, last_loss, nullptr);
}

int ves3d_hip_lpc_dataset_loss(void *h, double *loss, double *target_ms) {
  if (!h) return VES3D_INVALID_PARAMETER;
  LearnedPC *l = (LearnedPC *)h;
  if (l->stored() < 1) return VES3D_INVALID_PARAMETER;
  return l->dataset_loss(l->stored(), loss, target_ms, nullptr);
}

}  // extern "C"

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
int bicgstab(long n, apply_fn A, const void *actx, apply_fn P, const void *pctx, double *x,
             const double *b, int *max_iter, double *tol, int *status, hipStream_t st) {
  DBuf r, rt, p, v, s, t, ph, sh;
  for (DBuf *d : {&r, &rt, &p, &v, &s, &t, &ph, &sh}) V3D_CHECK(d->alloc(n));
  int err = 0;
  double normb = dev_norm(n, b, st, &err);
  V3D_CHECK(err);
  if (normb == 0.0) normb = 1.0;
  V3D_CHECK(A(actx, x, r));
  V3D_CHECK(dev_axpby(n, 1.0, b, -1.0, r, st));  // r = b - Ax
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

// ------------------------------------------------------------- the step
struct Stepper {
  ves3d_step_params prm;
  long nv = 0, ng = 0, nc = 0, n = 0;
  hipStream_t st = nullptr;
  ForcesG forces;
  StokesG stokes;
  DBuf X, U, sig;                  // state
  DBuf f, f2, Sf, ten, ug, sg, cf, rhs, xsol, yres, tmpD;
  DBuf bprec, tprec;               // diagonal-spectral preconditioner
  int2 *cinfo = nullptr;
  LearnedPC *lpc = nullptr;        // not owned
  bool collect = false;            // record (A e, e) pairs for the learned PC
  bool prepared = false;
  int matvecs = 0;
  int mode_precision = 0;

  ~Stepper() {
    if (cinfo) (void)hipFree(cinfo);
  }

  int init(const ves3d_step_params &p_) {
    prm = p_;
    const int p = prm.p;
    if (p < 2 || p > 64 || prm.nves < 1) return VES3D_INVALID_PARAMETER;
    if (prm.p_up <= 0) prm.p_up = 2 * p;
    if (prm.filter_freq <= 0) prm.filter_freq = 2 * p / 3;
    nv = (long)prm.nves;
    ng = sh_ngrid(p);
    nc = sh_ncoef(p);
    n = 4 * nv * nc;
    V3D_CHECK(forces.init(p, prm.p_up, prm.filter_freq, nv, prm.bending_modulus));
    if (prm.repul_dist == 0.0) prm.repul_dist = 5e-2;  // src/Parameters.cc:46
    V3D_CHECK(stokes.init(p, prm.p_up, nv, prm.precision, prm.near_off ? 0 : 1, prm.repul_dist,
                          prm.r_near));
    for (DBuf *b : {&X, &U, &f, &f2, &Sf, &ug}) V3D_CHECK(b->alloc(3 * nv * ng));
    for (DBuf *b : {&sig, &ten, &sg}) V3D_CHECK(b->alloc(nv * ng));
    V3D_CHECK(cf.alloc(3 * nv * nc));
    V3D_CHECK(rhs.alloc(n));
    V3D_CHECK(xsol.alloc(n));
    V3D_CHECK(yres.alloc(n));
    V3D_CHECK(tmpD.alloc(n));
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
    prepared = true;
    return VES3D_SUCCESS;
  }

  // ImplicitMatvecPhysical (src/InterfacialVelocity.cc:921-979), velocity form
  int matvec(const double *x, double *y) {
    if (!prepared) return VES3D_INVALID_PARAMETER;
    ++matvecs;
    V3D_CHECK(unpack(x, ug, sg));
    V3D_CHECK(dev_axpby(3 * nv * ng, prm.dt, ug, 0.0, f2, st));  // Du = dt u
    V3D_CHECK(forces.linear_bending(f2, f, st));
    V3D_CHECK(forces.tensile(sg, f2, st));
    V3D_CHECK(dev_axpby(3 * nv * ng, 1.0, f2, 1.0, f, st));
    V3D_CHECK(stokes.apply(f, Sf, st));
    V3D_CHECK(forces.S.div(Sf, ten, st));
    V3D_CHECK(dev_axpby(nv * ng, 0.0, ten, -1.0, ten, st));      // ten = -div
    V3D_CHECK(dev_axpby(3 * nv * ng, -1.0, Sf, 1.0, ug, st));    // vox = u - Sf
    return pack(ug, ten, y);
  }

  // AssembleRhsVel (src/InterfacialVelocity.cc:715-787): the explicit
  // traction jump plus the repulsion force (:733)
  int assemble_rhs(double *b) {
    if (!prepared) return VES3D_INVALID_PARAMETER;
    k_shear<<<gb(nv * ng), 256, 0, st>>>(X, nv, ng, prm.shear_rate, ug);
    V3D_CHECK(launch_status());
    V3D_CHECK(forces.bending(f, st));
    V3D_CHECK(stokes.apply(f, Sf, st, true));
    V3D_CHECK(dev_axpby(3 * nv * ng, 1.0, Sf, 1.0, ug, st));
    V3D_CHECK(forces.S.div(ug, ten, st));
    return pack(ug, ten, b);
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
    if (s->collect && s->lpc) {  // training pair: input A x, target x
      V3D_CHECK(s->precond_diag(y, s->tmpD));
      V3D_CHECK(s->lpc->record(s->tmpD, x, s->nv, s->prm.p, s->cinfo, s->st));
    }
    return VES3D_SUCCESS;
  }
  static int Pdiag_cb(const void *ctx, const double *x, double *y) {
    return ((Stepper *)ctx)->precond_diag(x, y);
  }
  static int Plearn_cb(const void *ctx, const double *x, double *y) {
    Stepper *s = (Stepper *)ctx;
    V3D_CHECK(s->precond_diag(x, y));
    return s->lpc->apply_inplace(y, s->nv, s->prm.p, s->cinfo, s->st);
  }

  int step(ves3d_step_stats *stats) {
    using clk = std::chrono::steady_clock;
    auto t0 = clk::now();
    stokes.timing_begin();
    matvecs = 0;
    V3D_CHECK(prepare());
    V3D_HIP(hipStreamSynchronize(st));
    auto t1 = clk::now();
    V3D_CHECK(assemble_rhs(rhs));
    V3D_CHECK(pack(U, sig, xsol));  // AssembleInitial (src/InterfacialVelocity.cc:873-919)
    V3D_HIP(hipStreamSynchronize(st));
    auto t2 = clk::now();
    int it = prm.max_iter;
    double tol = prm.tol;
    int status = 0;
    apply_fn P = nullptr;
    const bool learned = prm.precond == 2 && lpc && lpc->ready();
    if (prm.precond == 1 || (prm.precond == 2 && !learned)) P = Pdiag_cb;
    if (learned) P = Plearn_cb;
    V3D_CHECK(bicgstab(n, A_cb, this, P, this, xsol, rhs, &it, &tol, &status, st));
    // true residual of the returned solution (src/InterfacialVelocity.cc:181-188)
    double true_rel = 0;
    V3D_CHECK(true_residual(xsol, rhs, &true_rel));
    bool fell_back = false;
    if (learned && (status != 100 || !(true_rel <= prm.tol * 1.05))) {
      // the learned preconditioner did not deliver: unpreconditioned
      // BiCGStab restarts from the initial guess and its result stands
      fell_back = true;
      V3D_CHECK(pack(U, sig, xsol));
      it = prm.max_iter;
      tol = prm.tol;
      V3D_CHECK(bicgstab(n, A_cb, this, nullptr, nullptr, xsol, rhs, &it, &tol, &status, st));
      V3D_CHECK(true_residual(xsol, rhs, &true_rel));
    }
    auto t3 = clk::now();
    V3D_CHECK(unpack(xsol, U, sig));
    V3D_CHECK(dev_axpby(3 * nv * ng, prm.dt, U, 1.0, X, st));  // dx = dt u
    V3D_HIP(hipStreamSynchronize(st));
    auto t4 = clk::now();
    double far_ms = 0;
    int far_calls = 0;
    V3D_CHECK(stokes.far_time(&far_ms, &far_calls));
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
    }
    return VES3D_SUCCESS;
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

int ves3d_hip_stepper_precond(void *h, const double *x, double *y) {
  if (!h || !x || !y) return VES3D_INVALID_PARAMETER;
  Stepper *s = (Stepper *)h;
  V3D_CHECK(s->precond_diag(x, y));
  return hip_status(hipStreamSynchronize(s->st));
}

int ves3d_hip_stepper_step(void *h, ves3d_step_stats *stats) {
  if (!h) return VES3D_INVALID_PARAMETER;
  return ((Stepper *)h)->step(stats);
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
  l->set_ready(true);
  return VES3D_SUCCESS;
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

int ves3d_hip_lpc_adam(void *h, float lr, ves3d_stream_t stream) {
  if (!h) return VES3D_INVALID_PARAMETER;
  return ((LearnedPC *)h)->adam(lr, (hipStream_t)stream);
}

int ves3d_hip_lpc_train(void *h, int epochs, int batch, float lr, double *last_loss) {
  if (!h) return VES3D_INVALID_PARAMETER;
  return ((LearnedPC *)h)->train(epochs, batch, lr, last_loss, nullptr);
}

}  // extern "C"

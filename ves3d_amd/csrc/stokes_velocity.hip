// stokes_velocity.hip -- StokesVelocity<Real> (src/StokesVelocity.cc) as a
// C-ABI handle: the Stokes velocity of single- and double-layer densities
// on the vesicles' surfaces (singular self term + far field + near-singular
// correction) or at targets off the surfaces (SetTrgCoord), FP64.
#include <new>

#include "implicit.hpp"

using v3d::StokesG;

namespace {
struct StokesVelocity {
  StokesG g;
  bool ready = false;  // sources set
};
}  // namespace

extern "C" int ves3d_hip_stokes_create(int p, int p_up, size_t nves, double r_near,
                                       void **handle_host) {
  if (!handle_host || p < 2 || p > 64 || nves < 1 || p_up < 0) return VES3D_INVALID_PARAMETER;
  *handle_host = nullptr;
  StokesVelocity *s = new (std::nothrow) StokesVelocity();
  if (!s) return VES3D_MEMORY_ERROR;
  // FP64, near-singular correction on, no repulsion force
  int e = s->g.init(p, p_up ? p_up : 2 * p, (long)nves, 0, 1, -1.0, r_near, 0.0);
  if (e == VES3D_SUCCESS) e = s->g.enable_dl();
  if (e != VES3D_SUCCESS) {
    delete s;
    return e;
  }
  *handle_host = s;
  return VES3D_SUCCESS;
}

extern "C" int ves3d_hip_stokes_destroy(void *handle) {
  delete static_cast<StokesVelocity *>(handle);
  return VES3D_SUCCESS;
}

extern "C" int ves3d_hip_stokes_set_src(void *handle, const double *X, ves3d_stream_t stream) {
  if (!handle || !X) return VES3D_INVALID_PARAMETER;
  StokesVelocity *s = static_cast<StokesVelocity *>(handle);
  s->ready = false;
  V3D_CHECK(s->g.setup(X, (hipStream_t)stream));
  s->ready = true;
  return VES3D_SUCCESS;
}

extern "C" int ves3d_hip_stokes_eval(void *handle, const double *f_sl, const double *f_dl,
                                     const double *trg, size_t n_trg, double *vel,
                                     ves3d_stream_t stream) {
  if (!handle || !vel || (!f_sl && !f_dl)) return VES3D_INVALID_PARAMETER;
  StokesVelocity *s = static_cast<StokesVelocity *>(handle);
  if (!s->ready) return VES3D_INVALID_PARAMETER;
  hipStream_t st = (hipStream_t)stream;
  if (!trg) V3D_CHECK(s->g.apply(f_sl, vel, st, false, false, f_dl));
  else V3D_CHECK(s->g.eval_off(f_sl, f_dl, trg, (long)n_trg, vel, st));
  return v3d::hip_status(hipStreamSynchronize(st));
}

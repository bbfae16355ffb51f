"""Typed wrappers over the direct Stokes entry points (Device<>::DirectStokes,
DirectStokesDoubleLayer, StokesAlltoAll and the PVFMMEval stand-in).

Arguments are torch tensors on the GPU in the reference layouts; results are
written in place.  Each wrapper checks shapes and dtypes on the host before
the launch.
"""
import ctypes

import torch

from . import call, ptr, stream_ptr

_SUFFIX = {torch.float64: "f64", torch.float32: "f32"}


def _chk(t, dtype, numel=None, name="tensor"):
    if t is None:
        return
    if not t.is_cuda or t.dtype != dtype or not t.is_contiguous():
        raise ValueError("%s must be a contiguous %s CUDA tensor" % (name, dtype))
    if numel is not None and t.numel() < numel:
        raise ValueError("%s has %d elements, need %d" % (name, t.numel(), numel))


def direct_stokes(src, den, qw, src_stride, trg_stride, n_surfs, trg,
                  trg_idx_head, trg_idx_tail, pot, stream=None):
    """Device<DT>::DirectStokes (include/Device.h:205-209)."""
    dt = src.dtype
    for n, t in (("src", src), ("den", den), ("trg", trg), ("pot", pot)):
        _chk(t, dt, None, n)
    _chk(src, dt, 3 * src_stride * n_surfs, "src")
    _chk(den, dt, 3 * src_stride * n_surfs, "den")
    _chk(trg, dt, 3 * trg_stride * n_surfs, "trg")
    _chk(pot, dt, 3 * trg_stride * n_surfs, "pot")
    _chk(qw, dt, src_stride, "qw")
    call("ves3d_hip_direct_stokes_" + _SUFFIX[dt], ptr(src), ptr(den), ptr(qw),
         src_stride, trg_stride, n_surfs, ptr(trg), trg_idx_head, trg_idx_tail,
         ptr(pot), stream_ptr(stream))


def direct_stokes_dl(src, norm, den, qw, src_stride, trg_stride, n_surfs, trg,
                     trg_idx_head, trg_idx_tail, pot, stream=None):
    """Device<DT>::DirectStokesDoubleLayer (include/Device.h:211-215)."""
    dt = src.dtype
    for n, t, sz in (("src", src, src_stride), ("norm", norm, src_stride),
                     ("den", den, src_stride), ("trg", trg, trg_stride),
                     ("pot", pot, trg_stride)):
        _chk(t, dt, 3 * sz * n_surfs, n)
    _chk(qw, dt, src_stride, "qw")
    call("ves3d_hip_direct_stokes_dl_" + _SUFFIX[dt], ptr(src), ptr(norm), ptr(den),
         ptr(qw), src_stride, trg_stride, n_surfs, ptr(trg), trg_idx_head,
         trg_idx_tail, ptr(pot), stream_ptr(stream))


def stokes_all2all(src, den, pot, mixed=False, stream=None):
    """StokesAlltoAll (include/CPUKernels.h:46-49); point-major [np][3]."""
    n = src.numel() // 3
    for nm, t in (("src", src), ("den", den), ("pot", pot)):
        _chk(t, torch.float64, 3 * n, nm)
    call("ves3d_hip_stokes_all2all_" + ("f32c" if mixed else "f64"), ptr(src),
         ptr(den), n, ptr(pot), stream_ptr(stream))


def direct_eval(src_pos, sl_den, dl_den, trg_pos, trg_vel, src_per_group=0,
                trg_per_group=0, mixed=False, stream=None):
    """Direct-sum PVFMMEval stand-in (include/PVFMMInterface.h:53)."""
    ns = src_pos.numel() // 3
    nt = trg_pos.numel() // 3
    _chk(src_pos, torch.float64, 3 * ns, "src_pos")
    _chk(sl_den, torch.float64, 3 * ns, "sl_den")
    _chk(dl_den, torch.float64, 6 * ns, "dl_den")
    _chk(trg_pos, torch.float64, 3 * nt, "trg_pos")
    _chk(trg_vel, torch.float64, 3 * nt, "trg_vel")
    call("ves3d_hip_direct_eval_" + ("f32c" if mixed else "f64"), ptr(src_pos),
         ptr(sl_den), ptr(dl_den), ns, ptr(trg_pos), ptr(trg_vel), nt,
         src_per_group, trg_per_group, stream_ptr(stream))


class StokesVelocity:
    """StokesVelocity<Real> (src/StokesVelocity.cc) on the GPU, FP64: the
    velocity of single- and double-layer densities on the vesicles' surfaces
    or at targets off them (ves3d_hip_stokes_*)."""

    def __init__(self, nves, p, p_up=0, r_near=0.0):
        h = ctypes.c_void_p()
        call("ves3d_hip_stokes_create", p, p_up, nves, r_near, ctypes.byref(h))
        self.h, self.nves, self.p = h, nves, p
        self.ng = 2 * p * (p + 1)

    def close(self):
        if self.h:
            call("ves3d_hip_stokes_destroy", self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_src(self, X):
        _chk(X, torch.float64, 3 * self.nves * self.ng, "X")
        torch.cuda.synchronize()
        call("ves3d_hip_stokes_set_src", self.h, ptr(X), None)

    def __call__(self, f_sl=None, f_dl=None, trg=None):
        """Velocity [nves][3][ng] on the surfaces (trg None) or [n][3] at
        targets trg [n][3] off the surfaces."""
        n = 3 * self.nves * self.ng
        _chk(f_sl, torch.float64, n, "f_sl")
        _chk(f_dl, torch.float64, n, "f_dl")
        if trg is None:
            vel = torch.empty((self.nves, 3, self.ng), dtype=torch.float64, device="cuda")
            nt = 0
        else:
            nt = trg.numel() // 3
            _chk(trg, torch.float64, 3 * nt, "trg")
            vel = torch.empty((nt, 3), dtype=torch.float64, device="cuda")
        torch.cuda.synchronize()
        call("ves3d_hip_stokes_eval", self.h, ptr(f_sl), ptr(f_dl), ptr(trg), nt, ptr(vel), None)
        return vel

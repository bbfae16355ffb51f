"""Typed wrappers over the spherical-harmonic, rotation and singular
self-interaction entry points (SphericalHarmonics<Real> and the self term of
StokesVelocity).  Tensors are CUDA tensors in the reference layouts."""
import torch

from . import call, ptr, stream_ptr

_SUF = {torch.float64: "f64", torch.float32: "f32"}


def ncoef(p):
    return p * (p + 2)


def ngrid(p):
    return 2 * p * (p + 1)


def _chk(t, numel, dtype, name):
    if t is None:
        return
    if not t.is_cuda or not t.is_contiguous() or t.dtype != dtype:
        raise ValueError("%s must be a contiguous %s CUDA tensor" % (name, dtype))
    if t.numel() < numel:
        raise ValueError("%s has %d elements, need %d" % (name, t.numel(), numel))


def shc2grid(S, N, p0, p1, X, X_theta=None, X_phi=None, stream=None):
    dt = S.dtype
    _chk(S, N * ncoef(p0), dt, "S")
    for nm, t in (("X", X), ("X_theta", X_theta), ("X_phi", X_phi)):
        _chk(t, N * ngrid(p1), dt, nm)
    call("ves3d_hip_shc2grid_" + _SUF[dt], ptr(S), N, p0, p1, ptr(X), ptr(X_theta), ptr(X_phi),
         stream_ptr(stream))


def grid2shc(X, N, p0, p1, S, stream=None):
    dt = X.dtype
    _chk(X, N * ngrid(p0), dt, "X")
    _chk(S, N * ncoef(p1), dt, "S")
    call("ves3d_hip_grid2shc_" + _SUF[dt], ptr(X), N, p0, p1, ptr(S), stream_ptr(stream))


def shc2grid_transpose(X, N, p0, p1, S, stream=None):
    dt = X.dtype
    _chk(X, N * ngrid(p0), dt, "X")
    _chk(S, N * ncoef(p1), dt, "S")
    call("ves3d_hip_shc2grid_transpose_" + _SUF[dt], ptr(X), N, p0, p1, ptr(S), stream_ptr(stream))


def shc2pole(S, N, p0, P, stream=None):
    dt = S.dtype
    _chk(S, N * ncoef(p0), dt, "S")
    _chk(P, 2 * N, dt, "P")
    call("ves3d_hip_shc2pole_" + _SUF[dt], ptr(S), N, p0, ptr(P), stream_ptr(stream))


def rotate_all(S, N, p, dof, out, stream=None):
    _chk(S, N * dof * ncoef(p), torch.float64, "S")
    _chk(out, N * (p + 1) * p * dof * ncoef(p), torch.float64, "out")
    call("ves3d_hip_rotate_all_f64", ptr(S), N, p, dof, ptr(out), stream_ptr(stream))


def rotate_transpose(S, N, p, dof, out, stream=None):
    n = N * (p + 1) * p * dof * ncoef(p)
    _chk(S, n, torch.float64, "S")
    _chk(out, n, torch.float64, "out")
    call("ves3d_hip_rotate_transpose_f64", ptr(S), N, p, dof, ptr(out), stream_ptr(stream))


def singular_build(X0, nves, p0, p1, SL=None, DL=None, stream=None):
    """StokesSingularInteg: X0 [nves][3][ngrid(p0)] FP64 -> SL/DL [nves][3Nc][3Nc]."""
    _chk(X0, nves * 3 * ngrid(p0), torch.float64, "X0")
    m = SL if SL is not None else DL
    dt = m.dtype
    n = nves * (3 * ncoef(p0)) ** 2
    _chk(SL, n, dt, "SL")
    _chk(DL, n, dt, "DL")
    call("ves3d_hip_singular_build_" + _SUF[dt], ptr(X0), nves, p0, p1, ptr(SL), ptr(DL),
         stream_ptr(stream))


def self_apply(SL, F, DL, G, nves, p, V, stream=None):
    m = SL if SL is not None else DL
    dt = m.dtype
    n = 3 * ncoef(p)
    _chk(SL, nves * n * n, dt, "SL")
    _chk(DL, nves * n * n, dt, "DL")
    _chk(F, nves * n, torch.float64, "F")
    _chk(G, nves * n, torch.float64, "G")
    _chk(V, nves * n, torch.float64, "V")
    call("ves3d_hip_self_apply_" + _SUF[dt], ptr(SL), ptr(F), ptr(DL), ptr(G), nves, p, ptr(V),
         stream_ptr(stream))

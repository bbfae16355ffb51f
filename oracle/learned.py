"""FP64 numpy oracle of the learned preconditioner's solver interface.

TEST INFRASTRUCTURE ONLY: the checker for the canonical frames, the
gather / scatter between the Galerkin vector and the network rows, the
tension gauge and the training-pair recording of the device learned
preconditioner (ves3d_amd/csrc/learned_pc.hip, stepper.hip).  The learned
preconditioner has no counterpart in the reference; it fills the slot of
ImplicitPrecond (src/InterfacialVelocity.cc:1054-1109).

  frames(X, p)        per-vesicle orthogonal polar factor F of
                      C = sum_k w_k X(k) s(k)^T, s = (sin t cos p, sin t sin p,
                      cos t) the parameter sphere, w the Gauss-Legendre weights
  rows(imp, x)        the per-vesicle rows [nv][4 nc] of a Galerkin vector
                      (three velocity blocks, then tension; function-major
                      coefficient order inside each block)
  precond(imp, ...)   P(r)_v = z_v + s_v K F_v g_o(F_v^T z_v / s_v), z = D r,
                      s_v = |z_v|, g_o(z) = (g(z) - g(-z)) / 2 the odd part
                      of the network, K zeroing the tension coefficients of
                      degree > the divergence filter (src/Surface.cc:372)
  samples(imp, ...)   the (input, target) rows recorded from (A_loc e, e)
"""
import numpy as np

from . import mlp
from .ves import _degree_of_coeff, fm_to_om, om_to_fm


def frames(X, p):
    X = np.asarray(X, dtype=np.float64)
    x, w = np.polynomial.legendre.leggauss(p + 1)
    phi = np.pi * np.arange(2 * p) / p
    s = np.sqrt(1.0 - x * x)
    S = np.stack([np.outer(s, np.cos(phi)).ravel(), np.outer(s, np.sin(phi)).ravel(),
                  np.repeat(x, 2 * p)])                      # [3][ng]
    W = np.repeat(w, 2 * p)
    out = np.zeros((X.shape[0], 3, 3))
    for v in range(X.shape[0]):
        C = (X[v] * W[None, :]) @ S.T
        U, _, Vt = np.linalg.svd(C)
        out[v] = U @ Vt
    return out


def rows(imp, x):
    nv, nc, p = imp.nv, imp.nc, imp.p
    U = om_to_fm(x[:3 * nv * nc], p, 3 * nv).reshape(nv, 3 * nc)
    T = om_to_fm(x[3 * nv * nc:], p, nv)
    return np.concatenate([U, T], axis=1)


def unrows(imp, R):
    nv, nc, p = imp.nv, imp.nc, imp.p
    return np.concatenate([fm_to_om(R[:, :3 * nc].reshape(3 * nv, nc), p),
                           fm_to_om(R[:, 3 * nc:], p)])


def rotate_rows(R, F, nc, inverse):
    """Velocity blocks of every row by F_v^T (inverse=True: into the
    canonical frame) or F_v (back)."""
    out = R.copy()
    for v in range(R.shape[0]):
        Vb = R[v, :3 * nc].reshape(3, nc)
        out[v, :3 * nc] = ((F[v].T if inverse else F[v]) @ Vb).ravel()
    return out


def tension_gauge(p, tcut):
    """Mask [4 nc] of the row entries the correction may change."""
    nc = p * (p + 2)
    keep = np.ones(4 * nc, dtype=bool)
    keep[3 * nc:] = _degree_of_coeff(p) <= tcut
    return keep


def precond(imp, params, x, F, tcut=None, bf16_ops=False):
    """bf16_ops: the network's BF16-activation form (oracle/mlp.py)."""
    p, nc = imp.p, imp.nc
    tcut = 2 * p // 3 if tcut is None else tcut
    z = rows(imp, imp.precond(x))
    s = np.sqrt(np.sum(z * z, axis=1)) + 1e-300
    zc = rotate_rows(z, F, nc, True) / s[:, None]
    g = 0.5 * (mlp.forward(params, zc, bf16_ops)[0] - mlp.forward(params, -zc, bf16_ops)[0])
    g = rotate_rows(g, F, nc, False) * s[:, None]
    g[:, ~tension_gauge(p, tcut)] = 0.0
    return unrows(imp, z + g)


def samples(imp, e, F):
    """Rows recorded for the Krylov vector e: Z = F^T (D A_loc e)_v / s_v,
    T = F^T (e_v - (D A_loc e)_v) / s_v."""
    nc = imp.nc
    dr = rows(imp, imp.precond(imp.matvec(e, local=True)))
    er = rows(imp, e)
    s = np.sqrt(np.sum(dr * dr, axis=1)) + 1e-300
    zc = rotate_rows(dr, F, nc, True)
    tc = rotate_rows(er, F, nc, True)
    return zc / s[:, None], (tc - zc) / s[:, None]

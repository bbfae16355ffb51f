"""Pins the oracle's double-layer path and its targets off the surfaces
(oracle/ves.py StokesSurface with a DL density, eval_off; oracle/near.py
with trg_is_surf false; Implicit with a viscosity contrast) by closed forms
that a wrong sign, a missing jump or a wrong coefficient breaks:

  constant DL density c on a closed surface: -c inside, -c/2 on the
  surface (the grid's outward normal), 0 outside -- at the vesicle's own
  surface points (self term), at another vesicle's near surface points
  (near-singular correction of surface targets), and at targets off the
  surfaces inside and outside within r_near (the jump's sign flips for
  interior targets, src/NearSingular.cc:1458-1469);
  constant SL density c on a unit sphere: 2c/3 inside, the exterior flow of
  a sphere translating at U = 2c/3 outside;
  a rigid translation U of vesicles with viscosity contrasts lambda:
  vel_coeff U - D[dl_coeff U] = (1 + lambda)/2 U + (1 - lambda)/2 U = U, so
  the matvec maps (U, 0) to (U, 0) whatever lambda
  (src/InterfacialVelocity.cc:945-968).

Tolerances are the near-singular scheme's own accuracy at p = 12 with the
reference's r_near = 2 r_ves / sqrt(p_up): 1e-10 for the self terms and the
interior, 1e-5 at surface points near another vesicle, 1e-4 / 5e-4 off the
surfaces outside (the extrapolation of a varying exterior field).
"""
import numpy as np

from oracle import ves
from ves3d_amd import workloads as W

C = np.array([0.3, -1.1, 0.7])


def _close_pair(p, gap):
    import sys
    import os
    sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
    from test_step_forms_gpu import close_pair
    X = W.implicit_step_case(2, p, 0.05, seed=5)
    return close_pair(X, p, 0, 1, gap)


def test_constant_double_layer_on_the_surfaces():
    p = 12
    X = _close_pair(p, 0.3)
    ng = X.shape[2]
    st = ves.StokesSurface(X, p, 2 * p, near=True, want_dl=True)
    assert len(st.near.pairs) > 0
    d = np.zeros((2, 3, ng))
    d[0] = C[:, None]
    v = st(None, d=d)
    assert np.max(np.abs(v[0] + 0.5 * C[:, None])) < 1e-10
    assert np.max(np.abs(v[1])) < 1e-5


def test_constant_double_layer_off_the_surfaces():
    p = 12
    X = _close_pair(p, 0.3)
    ng = X.shape[2]
    st = ves.StokesSurface(X, p, 2 * p, near=True, want_dl=True)
    d = np.zeros((2, 3, ng))
    d[0] = C[:, None]
    n = ves.Surface(X, p).normal[0]
    idx = np.random.default_rng(1).choice(ng, 40, replace=False)
    for h in (0.02, 0.1, 0.3):
        To = np.ascontiguousarray((X[0][:, idx] + h * n[:, idx]).T)
        Ti = np.ascontiguousarray((X[0][:, idx] - h * n[:, idx]).T)
        vo = st.eval_off(To, None, d)
        assert len(st.near_off.pairs) > 0
        vi = st.eval_off(Ti, None, d)
        assert np.max(np.abs(vo)) < 1e-4, (h, np.max(np.abs(vo)))
        assert np.max(np.abs(vi + C)) < 1e-5, (h, np.max(np.abs(vi + C)))


def test_single_layer_off_a_sphere():
    p = 12
    Xs, _ = W.sphere_grid(p)
    cen = np.array([0.3, -0.2, 0.5])
    X = Xs[None] + cen[None, :, None]
    ng = Xs.shape[1]
    st = ves.StokesSurface(X, p, 2 * p, near=True)
    f = np.repeat(C[None, :, None], ng, axis=2)
    rng = np.random.default_rng(2)
    dirs = rng.standard_normal((40, 3))
    dirs /= np.linalg.norm(dirs, axis=1)[:, None]
    U = 2 * C / 3
    for r, tol in ((0.6, 1e-10), (0.95, 1e-8), (1.05, 5e-4), (1.4, 1e-6), (2.0, 1e-12)):
        T = np.ascontiguousarray(cen + r * dirs)
        v = st.eval_off(T, f, None)
        if r < 1:
            want = np.repeat(U[None], len(T), axis=0)
        else:
            x = T - cen
            rr = np.linalg.norm(x, axis=1)[:, None]
            ux = (x @ U)[:, None]
            want = 0.75 * (U / rr + ux * x / rr ** 3) + 0.25 * (U / rr ** 3 - 3 * ux * x / rr ** 5)
        assert np.max(np.abs(v - want)) < tol, (r, np.max(np.abs(v - want)))


def test_rigid_translation_with_viscosity_contrast():
    p = 12
    X = _close_pair(p, 0.3)
    ng = X.shape[2]
    U = np.random.default_rng(0).standard_normal((2, 3))
    u = np.repeat(U[:, :, None], ng, axis=2)
    for lam in ([5.0, 0.2], [1.0, 3.0]):
        imp = ves.Implicit(X, p, dt=1e-3, shear=0.0, viscosity_contrast=lam)
        assert imp.has_contrast
        x = imp.pack(u, np.zeros((2, ng)))
        y = imp.matvec(x)
        assert np.max(np.abs(y - x)) < 1e-5 * np.max(np.abs(x)), lam
    # a contrast changes the matvec of a deformation
    rng = np.random.default_rng(3)
    xr = rng.standard_normal(4 * 2 * p * (p + 2))
    y1 = ves.Implicit(X, p, dt=1e-3, shear=0.0).matvec(xr)
    y5 = ves.Implicit(X, p, dt=1e-3, shear=0.0, viscosity_contrast=5.0).matvec(xr)
    assert np.linalg.norm(y5 - y1) > 0.1 * np.linalg.norm(y1)

"""Pins the oracle of the learned preconditioner's solver interface
(oracle/learned.py) and the two facts its design rests on:

  frames             orthogonal, and equivariant: a shape rotated by R has
                     frame R F (so F^T x is the same for both);
  local operator     A_loc (ImplicitApply without far field and near
                     correction) equals every vesicle's operator alone, and
                     rotating a vesicle rotates its operator's velocity blocks:
                     A_loc(R X) (R e) = R (A_loc(X) e);
  recorded samples   identical for a suspension and the copy with every
                     vesicle rotated about its centre;
  tension gauge      the implicit operator never writes the tension
                     coefficients of degree > 2p/3 (Surface::div filters them,
                     src/Surface.cc:372) and they span its null space: a
                     preconditioner writing them moves the velocity of the
                     converged solve, the gauged one does not.
"""
import numpy as np

from oracle import learned as L
from oracle import ves
from ves3d_amd import workloads as W


def _rot(rng):
    return W.random_rotation(rng)


def _case(nv, p, seed, vf=0.05):
    return W.suspension(W.perturbed_shapes(nv, p, amp=0.05, seed=seed), vf, seed=seed)


def test_frames_orthogonal_and_equivariant():
    p = 8
    X = W.ellipsoid_shapes(3, p, seed=4)[0]
    rng = np.random.default_rng(1)
    F = L.frames(X, p)
    for v in range(3):
        assert np.allclose(F[v].T @ F[v], np.eye(3), atol=1e-13)
        R = _rot(rng)
        Fr = L.frames((R @ X[v])[None], p)[0]
        assert np.allclose(Fr, R @ F[v], atol=1e-12)
        # translation does not enter the degree-1 projection
        Ft = L.frames((X[v] + np.array([3.0, -1.0, 2.0])[:, None])[None], p)[0]
        assert np.allclose(Ft, F[v], atol=1e-12)


def test_local_operator_is_each_vesicle_alone():
    p = 6
    X = _case(3, p, seed=2)
    imp = ves.Implicit(X, p, dt=1e-3, shear=1.0, near=False)
    x = np.random.default_rng(0).standard_normal(4 * 3 * imp.nc)
    y = L.rows(imp, imp.matvec(x, local=True))
    xr = L.rows(imp, x)
    for v in range(3):
        one = ves.Implicit(X[v:v + 1], p, dt=1e-3, shear=1.0, near=False)
        yv = L.rows(one, one.matvec(L.unrows(one, xr[v:v + 1])))
        assert np.allclose(y[v], yv[0], atol=1e-12 * np.abs(yv).max())
    # the far field is what the full operator adds
    assert np.abs(L.rows(imp, imp.matvec(x)) - y).max() > 1e-3 * np.abs(y).max()


def test_local_operator_rotation_equivariance():
    p = 6
    X = W.ellipsoid_shapes(1, p, seed=6)[0]
    R = _rot(np.random.default_rng(3))
    a = ves.Implicit(X, p, dt=1e-3, shear=1.0, near=False)
    b = ves.Implicit((R @ X[0])[None], p, dt=1e-3, shear=1.0, near=False)
    e = L.rows(a, np.random.default_rng(4).standard_normal(4 * a.nc))
    Rs = R[None]
    ya = L.rows(a, a.matvec(L.unrows(a, e), local=True))
    yb = L.rows(b, b.matvec(L.unrows(b, L.rotate_rows(e, Rs, a.nc, False)), local=True))
    assert np.allclose(yb, L.rotate_rows(ya, Rs, a.nc, False), atol=1e-10 * np.abs(ya).max())


def test_samples_invariant_under_per_vesicle_rotation():
    p = 6
    nv = 3
    X = _case(nv, p, seed=8)
    rng = np.random.default_rng(9)
    Rs = np.stack([_rot(rng) for _ in range(nv)])
    c = X.mean(axis=2, keepdims=True)
    Xr = np.einsum("vij,vjk->vik", Rs, X - c) + c
    a = ves.Implicit(X, p, dt=1e-3, shear=1.0, near=False)
    b = ves.Implicit(Xr, p, dt=1e-3, shear=1.0, near=False)
    e = np.random.default_rng(10).standard_normal(4 * nv * a.nc)
    er = L.unrows(b, L.rotate_rows(L.rows(a, e), Rs, a.nc, False))
    Za, Ta = L.samples(a, e, L.frames(X, p))
    Zb, Tb = L.samples(b, er, L.frames(Xr, p))
    assert np.allclose(Za, Zb, atol=1e-10) and np.allclose(Ta, Tb, atol=1e-10)


def test_tension_gauge_is_the_null_space():
    p, nv = 6, 2
    X = _case(nv, p, seed=3)
    imp = ves.Implicit(X, p, dt=1e-3, shear=1.0, near=False)
    keep = np.tile(L.tension_gauge(p, 2 * p // 3), (nv, 1))
    rng = np.random.default_rng(0)
    y = L.rows(imp, imp.matvec(rng.standard_normal(4 * nv * imp.nc)))
    assert np.abs(y[~keep]).max() < 1e-13 * np.abs(y).max()
    b = imp.rhs()
    assert np.abs(L.rows(imp, b)[~keep]).max() < 1e-13 * np.abs(b).max()
    x0, _, _, st = ves.bicgstab(imp.matvec, b, np.zeros_like(b), 1e-11, 400, imp.precond)
    u0 = imp.unpack(x0)[0]
    n = b.size
    M = rng.standard_normal((n, n)) * 1e-3

    def noisy(v):
        return imp.precond(v) + M @ v

    def gauged(v):
        r = L.rows(imp, noisy(v))
        r[~keep] = L.rows(imp, imp.precond(v))[~keep]
        return L.unrows(imp, r)

    errs = []
    for P in (noisy, gauged):
        x, res, it, st = ves.bicgstab(imp.matvec, b, np.zeros_like(b), 1e-9, 400, P)
        assert st == "success"
        errs.append(np.linalg.norm(imp.unpack(x)[0] - u0) / np.linalg.norm(u0))
    assert errs[0] > 1e-4 and errs[1] < 1e-7, errs

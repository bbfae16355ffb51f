"""Pins the oracle of the implicit-step layer (oracle/ves.py).

  SHTrans matrices : the reference's own precomputed legTrans, legTransInv,
                     d1legTrans, d2legTrans files at p = 6, 12 (tests/golden).
  geometry         : unit sphere H = +1 (x_u x x_v points inward on this
                     grid), K = 1,
                     area 4 pi, volume 4 pi / 3; perturbed shape converges.
  forces           : bending force of a sphere vanishes; linearised mean
                     curvature of the position itself is the mean curvature.
  Stokes operator  : SL of a constant density on a lone unit sphere = (2/3) c.
  matvec           : exactly linear (test/InterfacialVelocityTest.cc:93-112).
  BiCGStab         : the diagonal system of test/BiCGStabTest.cc and random
                     nonsymmetric systems against numpy.linalg.solve.
"""
import os

import numpy as np
import pytest

from oracle import sph, ves
from ves3d_amd import workloads as W

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _file_blocks(name, p, inverse):
    """Unpacks a legTrans-style file into per-order [deg][lat] blocks."""
    v = np.loadtxt(os.path.join(GOLD, name))
    out = []
    pos = 0
    for m in range(p + 1):
        nd = p + 1 - m
        blk = v[pos:pos + nd * (p + 1)]
        pos += nd * (p + 1)
        if inverse:   # stored (lat x deg) column-major
            out.append(blk.reshape(nd, p + 1))
        else:         # stored (deg x lat) column-major
            out.append(blk.reshape(p + 1, nd).T)
    assert pos == v.size
    return out


@pytest.mark.parametrize("p", [6, 12])
def test_shtrans_matrices_match_reference_files(p):
    x, w = sph.legendre_nodes_weights(p)
    s2p = np.sqrt(2 * np.pi)
    P = sph._split_blocks(sph.legpoly(x, p), p)
    H = sph._split_blocks(sph.legpoly_deriv(x, p), p)
    Hl, Wl = sph.legpoly_ladder(x, p)
    Hl = sph._split_blocks(Hl, p)
    Wl = sph._split_blocks(Wl, p)
    fwd = _file_blocks("legTrans%d_double.txt" % p, p, False)
    inv = _file_blocks("legTransInv%d_double.txt" % p, p, True)
    d1 = _file_blocks("d1legTrans%d_double.txt" % p, p, True)
    d2 = _file_blocks("d2legTrans%d_double.txt" % p, p, True)
    for m in range(p + 1):
        assert np.allclose(fwd[m], s2p * P[m] * w[None, :], atol=1e-13), m
        assert np.allclose(inv[m], s2p * P[m], atol=1e-13), m
        # d1legTrans holds matlab/Ynm.m's Hn = -DPn, which is -LegPolyDeriv:
        # SHTrans's latitude derivative has the opposite sign of
        # SphericalHarmonics' (two independent derivations agree on it)
        assert np.allclose(d1[m], -s2p * H[m], atol=1e-12), m
        assert np.allclose(d1[m], s2p * Hl[m], atol=1e-12), m
        assert np.allclose(d2[m], s2p * Wl[m], atol=1e-11), m


@pytest.mark.parametrize("p", [6, 12, 16])
def test_sphere_geometry(p):
    X, _ = W.sphere_grid(p)
    S = ves.Surface(X[None], p)
    # reference orientation: outward x_u x x_v normal, H = -1 on the unit sphere
    assert np.allclose(S.normal[0], X, atol=1e-11)
    assert np.allclose(S.H, -1.0, atol=1e-11)
    assert np.allclose(S.K, 1.0, atol=1e-11)
    assert abs(S.area()[0] - 4 * np.pi) < 1e-11
    assert abs(S.volume()[0] - 4 * np.pi / 3) < 1e-11


def test_sphere_forces():
    p = 8
    X, _ = W.sphere_grid(p)
    S = ves.Surface(X[None] * 1.0, p)
    F = ves.Forces(S, 2 * p, 0.01)
    assert np.max(np.abs(F.bending())) < 1e-11
    h = S.linearized_mean_curv(S.X)
    assert np.allclose(h, S.H, atol=1e-12)
    # tension force of a uniform tension on a sphere: 2 sigma H n
    sig = np.full((1, X.shape[1]), 0.7)
    Fs = F.tensile(sig)
    assert np.allclose(Fs, 2 * 0.7 * -1.0 * S.normal, atol=1e-10)


def test_perturbed_geometry_converges():
    """Area of a band-limited perturbed shape is resolved spectrally."""
    Xs = W.perturbed_shapes(1, 6, amp=0.1, seed=2)
    a = []
    for q in (6, 12, 24):
        Xq = ves.resample_vec(Xs, 6, q)
        a.append(ves.Surface(Xq, q).area()[0])
    assert abs(a[1] - a[2]) < 1e-3 * abs(a[0] - a[2]) + 1e-12 or abs(a[1] - a[2]) < 1e-9


def test_stokes_surface_lone_sphere():
    p = 8
    X, _ = W.sphere_grid(p)
    st = ves.StokesSurface(X[None], p, 2 * p)
    c = np.array([0.3, -1.1, 0.7])
    f = np.broadcast_to(c[None, :, None], (1, 3, X.shape[1])).copy()
    assert np.allclose(st(f), 2.0 / 3.0 * f, atol=1e-9)


def _small_suspension(nv=2, p=6, seed=3):
    shapes = W.perturbed_shapes(nv, p, amp=0.05, seed=seed)
    return W.suspension(shapes, 0.05, seed=seed)


def test_matvec_is_linear():
    X = _small_suspension()
    imp = ves.Implicit(X, 6, dt=1e-3)
    n = 4 * imp.nv * imp.nc
    rng = np.random.default_rng(0)
    v1, v2 = rng.standard_normal(n), rng.standard_normal(n)
    assert np.max(np.abs(imp.matvec(np.zeros(n)))) < 1e-14
    a = imp.matvec(v1) + 2 * imp.matvec(v2)
    b = imp.matvec(v1 + 2 * v2)
    assert np.max(np.abs(a - b)) < 1e-10 * np.max(np.abs(a))


def test_bicgstab_diag():
    n = 4 * 12 * 14
    d = (np.arange(n) + 1.0) ** 2
    b = np.random.default_rng(1).uniform(size=n)
    # the reference test only prints the residual of this ill-conditioned
    # system (cond ~7e6); with the exact diagonal preconditioner it converges
    x, res, it, st = ves.bicgstab(lambda v: d * v, b, np.zeros(n), 1e-6, 1000, lambda v: v / d)
    assert st == "success" and res < 1e-6 and it <= 2
    assert np.linalg.norm(d * x - b) / np.linalg.norm(b) < 1e-6
    d2 = (np.arange(60) + 1.0) ** 2
    x, res, it, st = ves.bicgstab(lambda v: d2 * v, b[:60], np.zeros(60), 1e-6, 1000)
    assert st == "success" and np.linalg.norm(d2 * x - b[:60]) / np.linalg.norm(b[:60]) < 1e-6


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_bicgstab_random(seed):
    rng = np.random.default_rng(seed)
    n = 60
    A = np.eye(n) * 4 + rng.standard_normal((n, n)) * 0.3
    b = rng.standard_normal(n)
    x, res, it, st = ves.bicgstab(lambda v: A @ v, b, np.zeros(n), 1e-12, 500)
    assert st == "success"
    assert np.allclose(x, np.linalg.solve(A, b), atol=1e-9)
    P = np.diag(1.0 / np.diag(A))
    x2, res2, it2, st2 = ves.bicgstab(lambda v: A @ v, b, np.zeros(n), 1e-12, 500, lambda v: P @ v)
    assert st2 == "success" and np.allclose(x2, np.linalg.solve(A, b), atol=1e-9)


def test_implicit_solve_small():
    """One implicit solve on two vesicles at p=6 reaches the tolerance and its
    true residual confirms it (src/InterfacialVelocity.cc:181-188)."""
    X = _small_suspension()
    imp = ves.Implicit(X, 6, dt=1e-3, shear=1.0)
    b = imp.rhs()
    x, res, it, st = ves.bicgstab(imp.matvec, b, np.zeros_like(b), 1e-6, 200, imp.precond)
    assert st == "success"
    assert np.linalg.norm(imp.matvec(x) - b) / np.linalg.norm(b) < 1e-6


def test_reparam_is_tangential_and_lowers_the_energy():
    """InterfacialVelocity::reparam moves points along the surface: the
    normal component of the displacement is second order, the surface's area
    and volume barely change, and the degree-weighted energy decreases."""
    p = 8
    X = W.perturbed_shapes(1, p, amp=0.1, seed=12)
    Xr, it = ves.reparam(X, p)
    d = Xr - X
    S = ves.Surface(X, p)
    dn = np.abs(np.sum(d * S.normal, axis=1))
    assert it >= 1 and np.max(np.abs(d)) > 1e-4
    # tangential up to the band-limiting projection that follows each
    # tangential projection (it re-introduces a normal part)
    assert np.max(dn) < 0.2 * np.max(np.sqrt(np.sum(d * d, axis=1)))
    a0, v0 = ves.area_volume(X, p, 2 * p)
    a1, v1 = ves.area_volume(Xr, p, 2 * p)
    assert abs(a1[0] - a0[0]) / a0[0] < 1e-3 and abs(v1[0] - v0[0]) / v0[0] < 1e-3
    w = ves._rep_weights(p, 4)[ves._degree_of_coeff(p)]
    energy = lambda Y: float(np.sum(sph.grid2shc(Y.reshape(3, -1), p, p) ** 2 * w[None, :]))
    assert energy(Xr) < energy(X)


def test_area_volume_correction_restores_targets():
    p = 8
    X = W.perturbed_shapes(2, p, amp=0.1, seed=13)
    a0, v0 = ves.area_volume(X, p, 2 * p)
    # the filtered curvature and the resampling make the iteration converge
    # linearly (about x0.08 per step): from 1e-3 it needs ~9 of its 10 steps
    Xc, it = ves.area_volume_correction(X * 1.0005, p, 2 * p, a0, v0)
    a1, v1 = ves.area_volume(Xc, p, 2 * p)
    assert 1 <= it <= 10
    assert np.max(np.abs(a1 - a0) / a0) < 1e-11 and np.max(np.abs(v1 - v0) / v0) < 1e-11
    # already on target: no iteration, no change
    Xs, it0 = ves.area_volume_correction(X, p, 2 * p, a0, v0)
    assert it0 == 0 and np.array_equal(Xs, X)


def test_bending_precond_singular_degree():
    """dt n^3 = 1 (n = 10 at dt = 1e-3): the reference's 1/|1 - dt n^3| is
    +inf in IEEE arithmetic and falls back to 1.0
    (src/InterfacialVelocity.cc:680-682); neighbouring degrees keep the
    formula, and degrees whose value exceeds 1e3 in magnitude also give 1."""
    p = 12
    bp = ves.bending_precond(p, 1e-3)
    deg = ves._degree_of_coeff(p)
    assert np.all(bp[deg == 10] == 1.0)
    assert np.allclose(bp[deg == 9], 1.0 / abs(1.0 - 1e-3 * 729))
    assert np.allclose(bp[deg == 12], 1.0 / abs(1.0 - 1e-3 * 1728))
    assert np.all(bp[deg == 0] == 1.0)

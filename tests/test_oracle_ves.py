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


def test_surface_dumbbell_known_answers():
    """test/SurfaceTest.cc:66-176 on the p = 6 dumbbell (the first shape of
    precomputed/shape_gallery_6.txt), with that test's tolerances: area,
    volume, max |H|, the surface gradients and Laplacians of the coordinate
    functions, dot(Lap x, n), H - dot(Lap x, n)/2, H + div(n)/2, the
    linearised curvature of x itself, and n mapped to the tangent space."""
    p = 6
    X = np.load(os.path.join(GOLD, "dumbbell_6.npy"))[None]
    S = ves.Surface(X, p)
    assert abs(S.area()[0] / 16.2179377312307 - 1) < 5e-8
    assert abs(S.volume()[0] / 5.24886489292959 - 1) < 5e-8
    assert abs(np.max(np.abs(S.H)) / 1.376627062 - 1) < 8e-8
    want = [(1.01100423438481, 2.239995450856133), (1.011004234384816, 2.239995450856133),
            (1.004308085588217, 1.777133873450119)]
    lap = []
    for i, (g, l) in enumerate(want):
        gr = S.grad(X[:, i])
        lp = S.div(gr)
        lap.append(lp)
        assert abs(np.max(np.abs(gr)) / g - 1) < 1e-7, i
        assert abs(np.max(np.abs(lp)) / l - 1) < 1e-7, i
    hh = np.sum(np.stack(lap, axis=1) * S.normal, axis=1)
    assert abs(np.max(np.abs(hh)) / 2.428980517523748 - 1) < 1e-7
    assert abs(np.max(np.abs(S.H - 0.5 * hh)) / 0.348685011112687 - 1) < 1e-6
    assert abs(np.max(np.abs(S.H + 0.5 * S.div(S.normal))) / 0.2437253515 - 1) < 1e-6
    assert np.max(np.abs(S.linearized_mean_curv(S.X) - S.H)) < 1e-14
    n = S.normal
    assert np.max(np.abs(n - np.sum(n * S.normal, axis=1)[:, None] * S.normal)) < 1e-14


def _ellipsoid(p, seed=5):
    X = W.implicit_step_case(1, p, 0.2, seed=seed)
    return X - X.mean(axis=2, keepdims=True)


def _bending_energy(Y, p, q, kb):
    """E = 2 kb int H^2 dA on the surface resampled to order q (unfiltered
    curvature): the energy whose first variation the bending force is."""
    Yq = ves.resample_vec(Y, p, q)
    Sq = ves.Surface(Yq, q, filter_freq=q)
    qw = np.repeat(sph.legendre_weights(q), 2 * q)
    return 2 * kb * np.sum(Sq.H ** 2 * Sq.W * qw[None, :])


def test_bending_force_is_the_first_variation_of_the_energy():
    """On an ellipsoid (where both terms of the force are alive; on a sphere
    they vanish whatever their coefficients): moving the surface by
    eps phi n changes E = 2 kb int H^2 dA at the rate
    dE/deps = -2 int F_b . (phi n) dA, the sign making F_b the force that
    lowers the energy (bending_coeff = -kappa_b, src/VesicleProps.cc:41-43).
    A wrong coefficient of either term, or the opposite sign, fails by O(1).
    The linear bending force of the position itself is the bending force."""
    p, kb, eps = 16, 0.01, 1e-4
    X = _ellipsoid(p)
    S = ves.Surface(X, p)
    F = ves.Forces(S, 2 * p, kb)
    Fb = F.bending()
    phi = 0.1 * (X[0, 0] + 0.5 * X[0, 1] * X[0, 2])[None]
    d = phi[:, None] * S.normal
    dE = (_bending_energy(X + eps * d, p, 3 * p, kb) -
          _bending_energy(X - eps * d, p, 3 * p, kb)) / (2 * eps)
    qw = np.repeat(sph.legendre_weights(p), 2 * p)
    work = np.sum(np.sum(Fb * d, axis=1) * S.W * qw[None, :])
    assert abs(dE + 2 * work) < 1e-6 * abs(dE), (dE, work)
    assert np.max(np.abs(F.linear_bending(S.X) - Fb)) < 1e-12 * np.max(np.abs(Fb))


def test_tension_force_is_the_first_variation_of_the_area():
    """A uniform tension sigma: moving the surface by eps phi n changes the
    area at the rate dA/deps = -int F_s . (phi n) dA / sigma (F_s = 2 sigma H n
    + grad sigma, src/InterfacialForce.cc:109-112), on an ellipsoid."""
    p, eps, sig = 16, 1e-4, 0.7
    X = _ellipsoid(p)
    S = ves.Surface(X, p)
    F = ves.Forces(S, 2 * p, 0.01)
    Fs = F.tensile(np.full((1, X.shape[2]), sig))
    phi = 0.1 * (X[0, 0] + 0.5 * X[0, 1] * X[0, 2])[None]
    d = phi[:, None] * S.normal

    def area(Y):
        return ves.Surface(ves.resample_vec(Y, p, 3 * p), 3 * p).area()[0]

    dA = (area(X + eps * d) - area(X - eps * d)) / (2 * eps)
    qw = np.repeat(sph.legendre_weights(p), 2 * p)
    work = np.sum(np.sum(Fs * d, axis=1) * S.W * qw[None, :])
    assert abs(dA + work / sig) < 1e-6 * abs(dA), (dA, work)


def test_lone_vesicle_step_is_inextensible_and_dissipative():
    """A lone ellipsoid in fluid at rest: the step's velocity is surface-
    divergence free (the tension rows) and lowers the bending energy (the
    interfacial force does positive work on the fluid).  A sign slip
    between the velocity and tension rows, or in the bending force, fails
    one of the two."""
    p, kb, dt = 12, 0.01, 1e-2
    X = _ellipsoid(p)
    imp = ves.Implicit(X, p, dt=dt, kb=kb, shear=0.0)
    b = imp.rhs()
    x, res, it, st = ves.bicgstab(imp.matvec, b, np.zeros_like(b), 1e-10, 500, imp.precond)
    assert st == "success"
    u, _ = imp.unpack(x)
    S = ves.Surface(X, p)
    assert np.max(np.abs(S.div(u))) < 1e-8 * np.max(np.abs(u))
    h = 1e-4
    dEdt = (_bending_energy(X + h * u, p, 2 * p, kb) - _bending_energy(X - h * u, p, 2 * p, kb)) / (2 * h)
    assert dEdt < 0, dEdt


def test_gravity_force_total_is_weight():
    """InterfacialForce::gravityForce: the traction rho (g . x) n integrates
    over a closed surface to rho V g (divergence theorem), on an ellipsoid;
    zero without an excess density."""
    p = 16
    X = _ellipsoid(p) + np.array([0.3, -0.2, 1.5])[None, :, None]
    S = ves.Surface(X, p)
    g = np.array([0.2, -0.1, -1.0])
    F = ves.Forces(S, 2 * p, 0.01, excess_density=0.7, gravity=g)
    Fg = F.gravity(X)
    qw = np.repeat(sph.legendre_weights(p), 2 * p)
    tot = np.sum(Fg[0] * (S.W[0] * qw)[None, :], axis=1)
    assert np.allclose(tot, 0.7 * S.volume()[0] * g, atol=1e-10), (tot, 0.7 * S.volume()[0] * g)
    F0 = ves.Forces(S, 2 * p, 0.01)
    assert np.max(np.abs(F0.gravity(X))) == 0.0

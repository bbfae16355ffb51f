"""Accuracy of the two-level far field (oracle/ves.py StokesSurface with
far_coarse), both levels in the oracle: on the full-step workload's
ellipsoids (64 vesicles, p = 16, 20% volume fraction), a wave of 64 targets
that reads a vesicle's p-grid sources (17 x 32 points, the p-grid
quadrature) instead of its upsampled ones (33 x 64 + 2) where all its
targets are at least 3 radii from the vesicle's centre changes the far
field by less than 1e-6 of its size (L2 and max-norm), for the explicit
bending force and for a smooth force of the kind the solver's matvecs
carry.  The p-grid rule's aliasing error decays like (R / D)^p, so the
threshold is for p = 16; at p = 8 the same threshold leaves about 1e-5.

The flags themselves: a wave is flagged for a vesicle exactly when the
closest of its targets is at least kappa R from the centre, checked by
brute force.

The third level (far_coarse2): beyond kappa2 radii a wave reads the vesicle
at order p/2 -- the degree-p surface at the order-p/2 grid points, the
density f W / sin(theta) truncated to degree p/2 with the order-p/2
Gauss-Legendre weights.  Pins: the level's densities integrate f over the
surface exactly (the truncation keeps degree 0 and the rule is exact to
degree p + 1), and on a sphere they also reproduce the first moment
int (y - c) f^T dA (y is degree 1); with kappa = 3, kappa2 = 6 at p = 16
the three-level far field stays within 1e-6 of the upsampled one (L2 and
max) on the same workload.
"""
import numpy as np

from oracle import sph, ves
from ves3d_amd import workloads as W


def _smooth(nf, p, rng):
    """Random band-limited fields [nf][ng] whose degree-n coefficients have
    standard deviation (1 + n)^-1.5."""
    c = np.zeros((nf, p * (p + 2)))
    off, L = sph.block_offsets(p)
    for j in range(2 * p):
        m = (j + 1) // 2
        for k in range(L[j]):
            c[:, off[j] + k] = rng.standard_normal(nf) / (1 + m + k) ** 1.5
    return sph.shc2grid(c, p, p)


def test_two_level_far_field_accuracy_p16():
    nv, p, kappa = 64, 16, 3.0
    ng = 2 * p * (p + 1)
    X = W.implicit_step_case(nv, p, 0.2, seed=W.SEED_BASE + 3)
    st = ves.StokesSurface(X, p, 2 * p, near=False, far_coarse=kappa, self_op=False)
    fl = st.coarse_flags
    assert 0.5 < fl.mean() < 0.99, fl.mean()
    S = ves.Surface(X, p)
    F = ves.Forces(S, 2 * p, 0.01)
    rng = np.random.default_rng(0)
    u = _smooth(3 * nv, p, rng).reshape(nv, 3, ng)
    sg = _smooth(nv, p, rng)
    for name, f in (("bending", F.bending()), ("matvec", F.linear_bending(1e-3 * u) + F.tensile(sg))):
        v1 = st.far_field(f, False)
        v2 = st.far_field(f, True)
        l2 = np.linalg.norm(v2 - v1) / np.linalg.norm(v1)
        mx = np.max(np.abs(v2 - v1)) / np.max(np.abs(v1))
        print(name, "coarse fraction %.3f, L2 %.2e, max %.2e" % (fl.mean(), l2, mx))
        assert l2 < 1e-6 and mx < 1e-6, (name, l2, mx)


def test_coarse_flags_brute_force():
    nv, p, kappa = 12, 6, 2.5
    X = W.implicit_step_case(nv, p, 0.2, seed=7)
    st = ves.StokesSurface(X, p, 2 * p, near=False, far_coarse=kappa, self_op=False)
    mu2 = 2 * (2 * p) * (2 * p + 1) + 2
    src = st.scoord_far.reshape(nv, mu2, 3)
    c = src.mean(axis=1)
    R = np.sqrt(np.max(np.sum((src - c[:, None]) ** 2, axis=2), axis=1))
    T = st.trg
    for w in range(st.coarse_flags.shape[0]):
        for g in range(nv):
            dmin = min(np.linalg.norm(t - c[g]) for t in T[w * 64:(w + 1) * 64])
            assert st.coarse_flags[w, g] == (dmin >= kappa * R[g]), (w, g)


def test_three_level_far_field_accuracy_p16():
    nv, p = 64, 16
    ng = 2 * p * (p + 1)
    X = W.implicit_step_case(nv, p, 0.2, seed=W.SEED_BASE + 3)
    st = ves.StokesSurface(X, p, 2 * p, near=False, far_coarse=3.0, far_coarse2=6.0,
                           self_op=False)
    fl = st.coarse_flags
    assert 0.3 < np.mean(fl == 2) < 0.9 and np.mean(fl == 1) > 0.1, (np.mean(fl == 1),
                                                                    np.mean(fl == 2))
    S = ves.Surface(X, p)
    F = ves.Forces(S, 2 * p, 0.01)
    rng = np.random.default_rng(0)
    u = _smooth(3 * nv, p, rng).reshape(nv, 3, ng)
    sg = _smooth(nv, p, rng)
    for name, f in (("bending", F.bending()), ("matvec", F.linear_bending(1e-3 * u) + F.tensile(sg))):
        v1 = st.far_field(f, False)
        v3 = st.far_field(f, True)
        l2 = np.linalg.norm(v3 - v1) / np.linalg.norm(v1)
        mx = np.max(np.abs(v3 - v1)) / np.max(np.abs(v1))
        print(name, "level-2 fraction %.3f, L2 %.2e, max %.2e" % (np.mean(fl == 2), l2, mx))
        assert l2 < 1e-6 and mx < 1e-6, (name, l2, mx)


def test_level2_densities_integrate_the_force():
    p, nv = 12, 3
    ng = 2 * p * (p + 1)
    rng = np.random.default_rng(5)
    X = W.implicit_step_case(nv, p, 0.1, seed=11)
    st = ves.StokesSurface(X, p, 2 * p, near=False, far_coarse=2.0, far_coarse2=4.0,
                           self_op=False)
    f = _smooth(3 * nv, p, rng).reshape(nv, 3, ng)
    q2 = st.level2_density(f)                       # [nv][ng2][3]
    tot = np.sum(f * st.aqwp[:, None, :], axis=2)    # the p-grid rule: int f dA
    assert np.allclose(q2.sum(axis=1), tot, rtol=0, atol=1e-12 * np.abs(tot).max())
    # a sphere: the first moment too
    Xs, _ = W.sphere_grid(p)
    cen = np.array([0.4, -0.1, 0.2])
    Xs = (0.8 * Xs + cen[:, None])[None]
    sp = ves.StokesSurface(Xs, p, 2 * p, near=False, far_coarse=2.0, far_coarse2=4.0,
                           self_op=False)
    fs = _smooth(3, p, rng).reshape(1, 3, ng)
    q2 = sp.level2_density(fs)[0]
    y = sp.c2pos[0] - cen
    m1 = y.T @ q2
    Y = np.transpose(Xs[0], (1, 0)) - cen
    m1_ref = Y.T @ (fs[0] * sp.aqwp[0][None, :]).T
    assert np.allclose(q2.sum(axis=0), (fs[0] * sp.aqwp[0][None, :]).sum(axis=1), atol=1e-12)
    assert np.max(np.abs(m1 - m1_ref)) < 1e-12 * np.max(np.abs(m1_ref)), np.max(np.abs(m1 - m1_ref))


def test_three_level_flags_brute_force():
    nv, p, kappa, kappa2 = 12, 6, 2.0, 3.0
    X = W.implicit_step_case(nv, p, 0.2, seed=7)
    st = ves.StokesSurface(X, p, 2 * p, near=False, far_coarse=kappa, far_coarse2=kappa2,
                           self_op=False)
    mu2 = 2 * (2 * p) * (2 * p + 1) + 2
    src = st.scoord_far.reshape(nv, mu2, 3)
    c = src.mean(axis=1)
    R = np.sqrt(np.max(np.sum((src - c[:, None]) ** 2, axis=2), axis=1))
    T = st.trg
    fl = st.coarse_flags
    assert np.any(fl == 2) and np.any(fl == 1) and np.any(fl == 0)
    for w in range(fl.shape[0]):
        for g in range(nv):
            dmin = min(np.linalg.norm(t - c[g]) for t in T[w * 64:(w + 1) * 64])
            want = 2 if dmin >= kappa2 * R[g] else (1 if dmin >= kappa * R[g] else 0)
            assert fl[w, g] == want, (w, g)

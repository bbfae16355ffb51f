"""Pins of the MovePole oracle (oracle/movepole.py):

- all_rot_mats against the reference's precomputed/all_rot_mats_6_double.txt
  (latitudes 0, 1, 3 kept as tests/golden/all_rot_mats_6_lat013.npy) and
  sp_harm_rot_mats against SpHarmRotMats_p6_double.txt (all latitudes).  The
  reference files come from a MATLAB pipeline whose rotated grid lands on
  the poles up to rounding, so m >= 1 entries there carry errors of
  ~sqrt(eps); agreement is 2e-8;
- the Direct and ViaSpHarm schemes agree (test/MovePoleTest.cc:86 asks the
  squared L2 difference below 1e-12, i.e. the difference below 1e-6);
- a rotation moves the pole where the definition says: a band-limited field
  rotated to latitude i and longitude j equals the field evaluated at the
  rotated points; rotations compose to the identity at i = p (no move) up
  to the longitude reflection of the grid.
"""
import os

import numpy as np
import pytest

from oracle import movepole as MP
from oracle import sph

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
P = 6
N = MP.ngrid(P)


def test_all_rot_mats_match_reference_file():
    R = MP.all_rot_mats(P)
    ref = np.load(os.path.join(GOLD, "all_rot_mats_6_lat013.npy"))
    for a, i in enumerate((0, 1, 3)):
        assert np.max(np.abs(R[i] - ref[a])) < 2e-8, i


def test_sp_harm_rot_mats_match_reference_file():
    Rsh = MP.sp_harm_rot_mats(P)
    G = np.loadtxt(os.path.join(GOLD, "SpHarmRotMats_p6_double.txt"))
    pos = 0
    for i in range(P + 1):
        for deg, ln in enumerate(MP.degree_sizes(P)):
            blk = G[pos:pos + ln * ln].reshape(ln, ln).T
            pos += ln * ln
            assert np.max(np.abs(blk - Rsh[i][deg])) < 2e-8, (i, deg)
    assert pos == G.size


def test_sp_harm_blocks_preserve_l2_norm():
    """A rotation preserves the L2 norm within each degree: with the basis
    norms W = pi diag(2, 1, 1, ...) (the m = 0 function integrates cos^0 over
    the longitude, 2 pi, the others cos^2 / sin^2, pi), M^T W M = W."""
    for mats in MP.sp_harm_rot_mats(P):
        for deg, M in enumerate(mats[:-1]):          # degree p is truncated (2p of 2p+1)
            W = np.diag([2.0] + [1.0] * (M.shape[0] - 1))
            assert np.allclose(M.T @ W @ M, W, atol=5e-8), deg


def test_same_order_round_trip():
    ns = 3
    v = np.random.default_rng(1).standard_normal(ns * P * (P + 2))
    D = MP.collect_same_order(v, P, ns)
    assert np.array_equal(MP.collect_same_freq(D, P, ns), v)


def _dumbbell():
    return np.load(os.path.join(GOLD, "dumbbell_6.npy"))


def test_direct_vs_via_spharm_on_dumbbell():
    X = _dumbbell()                       # [3][84], the reference test's input
    err = 0.0
    for i in range(P + 1):
        for j in range(2 * P):
            a = MP.move_pole_direct(X, P, i, j)
            b = MP.move_pole_via_spharm(X, P, i, j)
            err = max(err, float(np.sum((a - b) ** 2)))
    assert err < 1e-12


def test_rotation_evaluates_at_rotated_points():
    rng = np.random.default_rng(3)
    S = rng.standard_normal((2, P * (P + 2)))
    X = sph.shc2grid(S, P, P)
    x = sph.legendre_nodes(P)
    lat = np.repeat(x, 2 * P)
    lon = np.tile(np.arange(2 * P) * np.pi / P, P + 1)
    s = np.sqrt(1 - lat * lat)
    pts = np.stack([s * np.cos(lon), s * np.sin(lon), lat])
    for i, j in ((0, 0), (2, 5), (4, 11)):
        out = MP.move_pole_direct(X, P, i, j)
        a = np.arccos(x[i]) - np.pi
        Ry = np.array([[np.cos(a), 0, np.sin(a)], [0, 1, 0], [-np.sin(a), 0, np.cos(a)]])
        Q = Ry @ pts
        phi = np.arctan2(Q[1], Q[0]) + j * np.pi / P     # longitude shift of the input field
        want = (MP.sh_basis(P, np.clip(Q[2], -1, 1), phi) @ S.T).T
        assert np.max(np.abs(out - want)) < 5e-8 * np.max(np.abs(want)), (i, j)


def test_aggregate_rotation_is_direct_for_every_longitude():
    X = _dumbbell()
    res = MP.aggregate_rotation([X, X[:1]], P, 2)
    for j in range(2 * P):
        assert np.array_equal(res[j][0], MP.move_pole_direct(X, P, 2, j))
        assert np.array_equal(res[j][1], MP.move_pole_direct(X[:1], P, 2, j))

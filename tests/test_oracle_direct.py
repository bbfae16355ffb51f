"""Pins the oracle's direct Stokes sums (oracle/direct.c).

1. Against the reference's own src/CPUKernels.cc, compiled into oracle/_ref
   by oracle/build_ref.sh (scalar and SSE kernels, single and double layer,
   all-to-all).
2. Against closed forms on the unit sphere, where the smooth Gauss-Legendre x
   uniform quadrature is exact (polynomial integrands):
     DL of a constant density f at any interior point = -f,
     DL of a constant density at an exterior point    =  0,
     SL of a constant density f at the centre          = (2/3) f.
"""
import ctypes

import numpy as np
import pytest

import oracle_lib as OL
from ves3d_amd import workloads as W

RNG = np.random.default_rng(7)


def _vesicle_case(nv, ss, ts):
    src = RNG.uniform(-1, 1, (nv, 3, ss))
    den = RNG.uniform(-1, 1, (nv, 3, ss))
    nrm = RNG.uniform(-1, 1, (nv, 3, ss))
    trg = RNG.uniform(-1, 1, (nv, 3, ts))
    qw = RNG.uniform(0.1, 1, ss)
    return src, den, nrm, trg, qw


def _relerr(a, b):
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


@pytest.mark.parametrize("use_qw", [True, False])
@pytest.mark.parametrize("nv,ss,ts,head,tail", [(3, 37, 29, 0, 29), (2, 64, 70, 5, 41), (1, 1, 1, 0, 1)])
def test_direct_stokes_matches_reference(orc, use_qw, nv, ss, ts, head, tail):
    ref = OL.load_ref()
    if ref is None:
        pytest.skip("oracle/_ref not built (reference sources absent)")
    src, den, _, trg, qw = _vesicle_case(nv, ss, ts)
    if ts == ss:
        trg[:, :, 0] = src[:, :, 0]  # a coincident pair exercises the r=0 rule
    q = qw if use_qw else None
    a = np.zeros((nv, 3, ts))
    b = np.zeros((nv, 3, ts))
    c = np.zeros((nv, 3, ts))
    args = (OL.ci(ss), OL.ci(ts), OL.ci(nv), OL.ci(head), OL.ci(tail), OL.p(q), OL.p(trg), OL.p(src), OL.p(den))
    orc.orc_direct_stokes(*args, OL.p(a))
    ref.ref_direct_stokes_f64(*args, OL.p(b))
    assert _relerr(a, b) < 1e-14
    if ss % 4 == 0:  # the SSE kernel requires it (src/CPUKernels.cc:610-611)
        # The reference's double SSE kernel takes 1/r from a float rsqrt plus one
        # float Newton step (src/CPUKernels.cc:719-729), so it is only
        # single-precision accurate; the oracle follows the scalar FP64 kernel.
        ref.ref_direct_stokes_sse_f64(*args, OL.p(c))
        assert 1e-12 < _relerr(a, c) < 2e-6
    assert np.all(a[:, :, :head] == 0) and np.all(a[:, :, tail:] == 0)


@pytest.mark.parametrize("use_qw", [True, False])
def test_direct_stokes_dl_matches_reference(orc, use_qw):
    ref = OL.load_ref()
    if ref is None:
        pytest.skip("oracle/_ref not built (reference sources absent)")
    nv, ss, ts = 3, 45, 33
    src, den, nrm, trg, qw = _vesicle_case(nv, ss, ts)
    q = qw if use_qw else None
    a = np.zeros((nv, 3, ts))
    b = np.zeros((nv, 3, ts))
    args = (OL.ci(ss), OL.ci(ts), OL.ci(nv), OL.ci(0), OL.ci(ts), OL.p(q), OL.p(trg), OL.p(src), OL.p(nrm), OL.p(den))
    orc.orc_direct_stokes_dl(*args, OL.p(a))
    ref.ref_direct_stokes_dl_f64(*args, OL.p(b))
    assert _relerr(a, b) < 1e-14


def test_all2all_matches_reference(orc):
    ref = OL.load_ref()
    if ref is None:
        pytest.skip("oracle/_ref not built (reference sources absent)")
    n = 301
    src = RNG.uniform(-2, 2, (n, 3))
    den = RNG.uniform(-1, 1, (n, 3))
    a = np.zeros((n, 3))
    b = np.zeros((n, 3))
    orc.orc_stokes_all2all(OL.p(src), OL.p(den), OL.cs(n), OL.p(a))
    ref.ref_stokes_all2all_f64(OL.p(src), OL.p(den), OL.cs(n), OL.p(b))
    assert _relerr(a, b) < 1e-14
    s32, d32 = src.astype(np.float32), den.astype(np.float32)
    a32 = np.zeros((n, 3), np.float32)
    b32 = np.zeros((n, 3), np.float32)
    orc.orc_stokes_all2all_f32(OL.p(s32), OL.p(d32), OL.cs(n), OL.p(a32))
    ref.ref_stokes_all2all_f32(OL.p(s32), OL.p(d32), OL.cs(n), OL.p(b32))
    assert _relerr(a32, b32) < 1e-5


def _sphere(p):
    X, qw = W.sphere_grid(p)
    return np.ascontiguousarray(X[None]), np.ascontiguousarray(qw)


def test_closed_forms_on_sphere(orc):
    p = 10
    X, qw = _sphere(p)
    npts = X.shape[2]
    f = np.array([0.3, -1.1, 0.7])
    den = np.ascontiguousarray(np.broadcast_to(f[None, :, None], (1, 3, npts)))
    nrm = X.copy()
    trg = np.zeros((1, 3, 3))
    trg[0, :, 1] = [0.2, -0.3, 0.1]       # interior, off-centre
    trg[0, :, 2] = [2.0, 1.5, -1.0]       # exterior
    pot = np.zeros((1, 3, 3))
    orc.orc_direct_stokes_dl(OL.ci(npts), OL.ci(3), OL.ci(1), OL.ci(0), OL.ci(3), OL.p(qw), OL.p(trg),
                             OL.p(X), OL.p(nrm), OL.p(den), OL.p(pot))
    assert np.allclose(pot[0, :, 0], -f, atol=1e-13)
    assert np.allclose(pot[0, :, 1], -f, atol=1e-6)   # off-centre: spectrally small error at p=8
    assert np.allclose(pot[0, :, 2], 0, atol=1e-6)
    sl = np.zeros((1, 3, 3))
    orc.orc_direct_stokes(OL.ci(npts), OL.ci(3), OL.ci(1), OL.ci(0), OL.ci(1), OL.p(qw), OL.p(trg),
                          OL.p(X), OL.p(den), OL.p(sl))
    assert np.allclose(sl[0, :, 0], 2.0 / 3.0 * f, atol=1e-13)


def test_eval_equals_device_kernels(orc):
    """The PVFMMEval stand-in (d = trg - src, +3/4pi) and the Device kernels
    (r = src - trg, -3/4pi) are the same operators."""
    p = 6
    X, qw = _sphere(p)
    npts = X.shape[2]
    f = RNG.uniform(-1, 1, (1, 3, npts))
    g = RNG.uniform(-1, 1, (1, 3, npts))
    trg = RNG.uniform(-0.5, 0.5, (1, 3, 11))
    sl_dev = np.zeros((1, 3, 11))
    dl_dev = np.zeros((1, 3, 11))
    orc.orc_direct_stokes(OL.ci(npts), OL.ci(11), OL.ci(1), OL.ci(0), OL.ci(11), OL.p(qw), OL.p(trg),
                          OL.p(X), OL.p(f), OL.p(sl_dev))
    orc.orc_direct_stokes_dl(OL.ci(npts), OL.ci(11), OL.ci(1), OL.ci(0), OL.ci(11), OL.p(qw), OL.p(trg),
                             OL.p(X), OL.p(X), OL.p(g), OL.p(dl_dev))
    spos = np.ascontiguousarray(X[0].T)
    sld = np.ascontiguousarray((f[0] * qw).T)
    dld = np.ascontiguousarray(np.concatenate([(g[0] * qw).T, X[0].T], axis=1))
    tpos = np.ascontiguousarray(trg[0].T)
    v = np.zeros((11, 3))
    orc.orc_stokes_sl_dl_eval(OL.p(spos), OL.p(sld), OL.p(dld), OL.cs(npts), OL.p(tpos), OL.p(v), OL.cs(11))
    assert _relerr(v.T, (sl_dev + dl_dev)[0]) < 1e-13
    # a coincident target is masked (contributes nothing), like r=0 in the Device kernel
    v2 = np.zeros((1, 3))
    orc.orc_stokes_sl_dl_eval(OL.p(spos), OL.p(sld), None, OL.cs(npts), OL.p(np.ascontiguousarray(spos[:1])),
                              OL.p(v2), OL.cs(1))
    assert np.all(np.isfinite(v2))

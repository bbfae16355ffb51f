"""GPU parity of the direct Stokes sums against the FP64 oracle.

Tolerances (max-norm relative error over the output):
  FP64 kernels           1e-12
  FP32 Device kernels    2e-5   (the reference's single-precision build)
  f32c (FP32 arithmetic, compensated accumulation, FP64 I/O)
                         5e-6 on vesicle surfaces, 2e-5 on random point clouds
                         (near-coincident random pairs make single terms large,
                         each carrying FP32 rounding)
"""
import numpy as np
import pytest
import torch

import oracle_lib as OL
from ves3d_amd import stokes as K
from ves3d_amd import workloads as W

pytestmark = pytest.mark.gpu
RNG = np.random.default_rng(11)


def _cuda(a, dtype=torch.float64):
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda", dtype)


def _relerr(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def _vesicle_case(nv, ss, ts, clustered=False):
    src = RNG.uniform(-1, 1, (nv, 3, ss))
    trg = RNG.uniform(-1, 1, (nv, 3, ts))
    if clustered:
        trg[:, :, : min(ss, ts)] = src[:, :, : min(ss, ts)]  # coincident pairs (r = 0)
    den = RNG.uniform(-1, 1, (nv, 3, ss))
    nrm = RNG.uniform(-1, 1, (nv, 3, ss))
    qw = RNG.uniform(0.1, 1, ss)
    return src, den, nrm, trg, qw


@pytest.mark.parametrize("nv,ss,ts,head,tail", [
    (2, 544, 544, 0, 544), (3, 37, 29, 3, 29), (1, 1, 1, 0, 1), (5, 312, 312, 100, 101),
    (4, 2114, 544, 0, 544)])
@pytest.mark.parametrize("use_qw", [True, False])
def test_direct_stokes_f64(hip, orc, nv, ss, ts, head, tail, use_qw):
    src, den, _, trg, qw = _vesicle_case(nv, ss, ts, clustered=True)
    q = qw if use_qw else None
    ref = np.zeros((nv, 3, ts))
    orc.orc_direct_stokes(OL.ci(ss), OL.ci(ts), OL.ci(nv), OL.ci(head), OL.ci(tail), OL.p(q), OL.p(trg),
                          OL.p(src), OL.p(den), OL.p(ref))
    pot = torch.zeros((nv, 3, ts), dtype=torch.float64, device="cuda")
    K.direct_stokes(_cuda(src), _cuda(den), _cuda(q) if use_qw else None, ss, ts, nv, _cuda(trg),
                    head, tail, pot)
    out = pot.cpu().numpy()
    assert _relerr(out, ref) < 1e-12
    assert np.all(out[:, :, :head] == 0) and np.all(out[:, :, tail:] == 0)


@pytest.mark.parametrize("use_qw", [True, False])
def test_direct_stokes_f32(hip, orc, use_qw):
    nv, ss, ts = 3, 544, 500
    src, den, _, trg, qw = _vesicle_case(nv, ss, ts)
    q = qw if use_qw else None
    ref = np.zeros((nv, 3, ts))
    orc.orc_direct_stokes(OL.ci(ss), OL.ci(ts), OL.ci(nv), OL.ci(0), OL.ci(ts), OL.p(q), OL.p(trg),
                          OL.p(src), OL.p(den), OL.p(ref))
    f = torch.float32
    pot = torch.zeros((nv, 3, ts), dtype=f, device="cuda")
    K.direct_stokes(_cuda(src, f), _cuda(den, f), _cuda(q, f) if use_qw else None, ss, ts, nv,
                    _cuda(trg, f), 0, ts, pot)
    assert _relerr(pot.double().cpu().numpy(), ref) < 2e-5


@pytest.mark.parametrize("dtype,tol", [(torch.float64, 1e-12), (torch.float32, 2e-5)])
@pytest.mark.parametrize("use_qw", [True, False])
def test_direct_stokes_dl(hip, orc, dtype, tol, use_qw):
    nv, ss, ts = 3, 544, 531
    src, den, nrm, trg, qw = _vesicle_case(nv, ss, ts)
    q = qw if use_qw else None
    ref = np.zeros((nv, 3, ts))
    orc.orc_direct_stokes_dl(OL.ci(ss), OL.ci(ts), OL.ci(nv), OL.ci(0), OL.ci(ts), OL.p(q), OL.p(trg),
                             OL.p(src), OL.p(nrm), OL.p(den), OL.p(ref))
    pot = torch.zeros((nv, 3, ts), dtype=dtype, device="cuda")
    K.direct_stokes_dl(_cuda(src, dtype), _cuda(nrm, dtype), _cuda(den, dtype),
                       _cuda(q, dtype) if use_qw else None, ss, ts, nv, _cuda(trg, dtype), 0, ts, pot)
    assert _relerr(pot.double().cpu().numpy(), ref) < tol


def test_direct_stokes_bad_range(hip):
    src = torch.zeros(3 * 8, dtype=torch.float64, device="cuda")
    pot = torch.zeros(3 * 8, dtype=torch.float64, device="cuda")
    import ves3d_amd
    with pytest.raises(ves3d_amd.Ves3dError) as e:
        K.direct_stokes(src, src, None, 8, 8, 1, src, 5, 9, pot)
    assert e.value.code == 10


@pytest.mark.parametrize("n", [1, 2, 257, 3001])
def test_all2all(hip, orc, n):
    src = RNG.uniform(-3, 3, (n, 3))
    if n > 2:
        src[1] = src[0]  # duplicate point: its pair is skipped (r = 0)
    den = RNG.uniform(-1, 1, (n, 3))
    ref = np.zeros((n, 3))
    orc.orc_stokes_all2all(OL.p(src), OL.p(den), OL.cs(n), OL.p(ref))
    for mixed, tol in ((False, 1e-12), (True, 2e-5)):
        pot = torch.full((n, 3), np.nan, dtype=torch.float64, device="cuda")
        K.stokes_all2all(_cuda(src), _cuda(den), pot, mixed=mixed)
        out = pot.cpu().numpy()
        assert np.all(np.isfinite(out))
        assert _relerr(out, ref) < tol, (mixed, _relerr(out, ref))


def test_all2all_empty(hip):
    z = torch.zeros((0, 3), dtype=torch.float64, device="cuda")
    K.stokes_all2all(z, z, z)


@pytest.mark.parametrize("sl,dl", [(True, False), (False, True), (True, True)])
@pytest.mark.parametrize("grouped", [False, True])
def test_direct_eval(hip, orc, sl, dl, grouped):
    nv, p = 6, 8
    pos, sld, dld, npv = W.stokes_direct_sum_case(nv=nv, p=p, volfrac=0.1, seed=5)
    ns = pos.shape[0]
    trg = pos.copy()  # targets are the surface points themselves
    spg = npv if grouped else 0
    a_sl = sld if sl else None
    a_dl = dld if dl else None
    ref = np.zeros((ns, 3))
    orc.orc_far_field(OL.p(pos), OL.p(a_sl), OL.p(a_dl), OL.cs(ns), OL.p(trg), OL.p(ref), OL.cs(ns),
                      OL.cs(spg), OL.cs(spg))
    for mixed, tol in ((False, 1e-12), (True, 5e-6)):
        out = torch.zeros((ns, 3), dtype=torch.float64, device="cuda")
        K.direct_eval(_cuda(pos), _cuda(a_sl) if sl else None, _cuda(a_dl) if dl else None, _cuda(trg), out,
                      spg, spg, mixed=mixed)
        e = _relerr(out.cpu().numpy(), ref)
        assert e < tol, (mixed, e)


def test_direct_eval_separate_targets(hip, orc):
    """Targets off the surfaces (n_trg != n_src) with the 128*eps mask."""
    pos, sld, dld, npv = W.stokes_direct_sum_case(nv=4, p=6, volfrac=0.1, seed=9)
    trg = RNG.uniform(pos.min(), pos.max(), (777, 3))
    trg[0] = pos[3]
    ref = np.zeros((777, 3))
    orc.orc_far_field(OL.p(pos), OL.p(sld), OL.p(dld), OL.cs(pos.shape[0]), OL.p(trg), OL.p(ref),
                      OL.cs(777), OL.cs(0), OL.cs(0))
    out = torch.zeros((777, 3), dtype=torch.float64, device="cuda")
    K.direct_eval(_cuda(pos), _cuda(sld), _cuda(dld), _cuda(trg), out)
    assert _relerr(out.cpu().numpy(), ref) < 1e-12


def test_full_size_far_field_properties(hip, orc):
    """512 vesicles at p=16 (278,528 points): f32c vs f64 on every target,
    and f64 vs the oracle on a sample of 96 targets."""
    pos, sld, dld, npv = W.stokes_direct_sum_case(nv=512, p=16)
    n = pos.shape[0]
    gp, gs, gd = _cuda(pos), _cuda(sld), _cuda(dld)
    v64 = torch.zeros((n, 3), dtype=torch.float64, device="cuda")
    v32 = torch.zeros((n, 3), dtype=torch.float64, device="cuda")
    K.direct_eval(gp, gs, gd, gp, v64, npv, npv)
    K.direct_eval(gp, gs, gd, gp, v32, npv, npv, mixed=True)
    a64 = v64.cpu().numpy()
    a32 = v32.cpu().numpy()
    rel = np.linalg.norm(a32 - a64) / np.linalg.norm(a64)
    assert rel < 1e-6, rel
    idx = np.sort(RNG.choice(n, 96, replace=False))
    # oracle restricted to the sampled targets: same groups (vesicle = idx // npv)
    ref = np.zeros((96, 3))
    for k, t in enumerate(idx):
        r1 = np.zeros((1, 3))
        orc.orc_far_field(OL.p(pos), OL.p(sld), OL.p(dld), OL.cs(n), OL.p(np.ascontiguousarray(pos[t:t + 1])),
                          OL.p(r1), OL.cs(1), OL.cs(n), OL.cs(0))
        own = slice((t // npv) * npv, (t // npv + 1) * npv)
        r_own = np.zeros((1, 3))
        orc.orc_far_field(OL.p(np.ascontiguousarray(pos[own])), OL.p(np.ascontiguousarray(sld[own])),
                          OL.p(np.ascontiguousarray(dld[own])), OL.cs(npv),
                          OL.p(np.ascontiguousarray(pos[t:t + 1])), OL.p(r_own), OL.cs(1), OL.cs(0), OL.cs(0))
        ref[k] = (r1 - r_own)[0]
    assert _relerr(a64[idx], ref) < 1e-10

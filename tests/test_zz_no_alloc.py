"""The allocation-free far-field evaluation (ves3d_hip_direct_eval_f64_noalloc)
against the allocating entry point it mirrors (ves3d_hip_direct_eval_f64),
called through the C ABI, twice in a row with the second input smaller (the
fixed scratch keeps the first call's contents), plus the size limit.

Bars: bit-identical results where both entry points use the same number of
source slices (every case here: the partial sums of 278,528 targets in 8
slices fit the 64 MiB scratch); the smaller case also against the oracle
to 1e-12.
"""
import numpy as np
import pytest
import torch

import oracle_lib as OL
from ves3d_amd import workloads as W

pytestmark = pytest.mark.gpu

MAX_TRG = (64 << 20) // 24


def _cu(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def _ptr(t):
    return None if t is None else t.data_ptr()


def _eval(lib, name, pos, sld, dld, trg, spg):
    ns, nt = pos.shape[0], trg.shape[0]
    out = torch.full((nt, 3), float("nan"), dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    rc = getattr(lib, name)(_ptr(pos), _ptr(sld), _ptr(dld), ns, _ptr(trg), _ptr(out), nt, spg,
                            spg, None)
    torch.cuda.synchronize()
    assert rc == 0, (name, rc)
    return out


def test_noalloc_matches_allocating_entry(hip, orc):
    lib = hip.load()
    cases = [(512, 16, True, True), (6, 8, True, False)]  # (nv, p, SL, DL); the second smaller
    for k, (nv, p, sl, dl) in enumerate(cases):
        pos, sld, dld, npv = W.stokes_direct_sum_case(nv=nv, p=p, volfrac=0.15 if nv > 6 else 0.1,
                                                      seed=5 + k)
        print("no-alloc far field: call %d, %d vesicles, p=%d, %d targets" %
              (k, nv, p, pos.shape[0]), flush=True)
        gp = _cu(pos)
        gs = _cu(sld) if sl else None
        gd = _cu(dld) if dl else None
        a = _eval(lib, "ves3d_hip_direct_eval_f64", gp, gs, gd, gp, npv)
        b = _eval(lib, "ves3d_hip_direct_eval_f64_noalloc", gp, gs, gd, gp, npv)
        assert torch.equal(a, b), float((a - b).abs().max())
        if nv <= 6:
            ref = np.zeros((pos.shape[0], 3))
            orc.orc_far_field(OL.p(pos), OL.p(sld if sl else None), OL.p(dld if dl else None),
                              OL.cs(pos.shape[0]), OL.p(pos), OL.p(ref), OL.cs(pos.shape[0]),
                              OL.cs(npv), OL.cs(npv))
            e = float(np.max(np.abs(b.cpu().numpy() - ref)) / np.max(np.abs(ref)))
            assert e < 1e-12, e


def test_noalloc_rejects_sizes_beyond_its_scratch(hip):
    lib = hip.load()
    # checked on the host before anything touches the (absent) buffers
    rc = lib.ves3d_hip_direct_eval_f64_noalloc(None, None, None, 1, None, None, MAX_TRG + 1, 0, 0,
                                              None)
    assert rc == 10  # VES3D_INVALID_PARAMETER

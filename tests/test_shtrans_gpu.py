"""GPU parity of the SHTrans-layout device SHT (forward, backward and its
derivatives, ScaleFreq, lowPassFilter, Resample) against the oracle
(oracle/ves.py, whose matrices are pinned by the reference's legTrans
fixtures).  Tolerance 1e-12 relative (FP64)."""
import numpy as np
import pytest
import torch

from oracle import sph, ves
from ves3d_amd import call, ptr, stream_ptr

pytestmark = pytest.mark.gpu
RNG = np.random.default_rng(31)


def cu(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


@pytest.mark.parametrize("p,nf", [(6, 5), (12, 3), (16, 7), (24, 3)])
def test_forward_backward(hip, p, nf):
    ng, nc = 2 * p * (p + 1), p * (p + 2)
    S = RNG.standard_normal((nf, nc))
    X = sph.shc2grid(S, p, p)
    out = torch.zeros(nf * nc, dtype=torch.float64, device="cuda")
    call("ves3d_hip_shtrans_forward", p, nf, ptr(cu(X)), ptr(out), stream_ptr())
    assert rel(out.cpu().numpy(), ves.fm_to_om(S, p)) < 1e-12
    om = cu(ves.fm_to_om(S, p))
    sht = ves.Sht(p)
    for du, dv in [(0, 0), (1, 0), (0, 1), (2, 0), (0, 2), (1, 1)]:
        g = torch.zeros(nf * ng, dtype=torch.float64, device="cuda")
        call("ves3d_hip_shtrans_backward", p, nf, ptr(om), du, dv, ptr(g), stream_ptr())
        ref = sht.backward(S, du, dv)
        assert rel(g.cpu().numpy().reshape(nf, ng), ref) < 1e-12, (du, dv)


def test_scale_freq_low_pass_resample(hip):
    p, nf = 12, 4
    ng, nc = 2 * p * (p + 1), p * (p + 2)
    S = RNG.standard_normal((nf, nc))
    scale = RNG.uniform(0.5, 2, nc)
    out = torch.zeros(nf * nc, dtype=torch.float64, device="cuda")
    g_in, g_sc = cu(ves.fm_to_om(S, p)), cu(scale)     # referenced until the kernel ran
    call("ves3d_hip_shtrans_scale_freq", p, nf, ptr(g_in), ptr(g_sc), ptr(out), stream_ptr())
    assert rel(out.cpu().numpy(), ves.fm_to_om(S * scale[None, :], p)) < 1e-14
    X = sph.shc2grid(S, p, p)
    g = torch.zeros(nf * ng, dtype=torch.float64, device="cuda")
    call("ves3d_hip_shtrans_low_pass", p, nf, 8, ptr(cu(X)), ptr(g), stream_ptr())
    assert rel(g.cpu().numpy().reshape(nf, ng), ves.Sht(p, 8).low_pass(X)) < 1e-12
    for q in (6, 24):
        r = torch.zeros(nf * 2 * q * (q + 1), dtype=torch.float64, device="cuda")
        call("ves3d_hip_resample", p, q, nf, ptr(cu(X)), ptr(r), stream_ptr())
        assert rel(r.cpu().numpy().reshape(nf, -1), ves.resample(X, p, q)) < 1e-12

"""The two- and three-level far field of the step (ves3d_step_params.
far_coarse, far_coarse2) against the oracle's restatement of the same scheme
(oracle/ves.py StokesSurface with far_coarse, far_coarse2): a wave of 64
targets reads a vesicle's p-grid sources instead of its upsampled ones
where all its targets are at least far_coarse times the vesicle's radius
from the vesicle's centre, and its order-p/2 sources beyond far_coarse2.

Bars (relative, max-norm unless stated):
  FP64 right-hand side / matvec      5e-9 (2e-8 with near-singular pairs; the
                                     host-built MatRotate's entries fed by the
                                     rotated pole differ from the oracle's by
                                     1e-8, which 24 vesicles at p = 6 show)
  precision 2 (FP32-compensated)     2e-5 L2
  pairs per evaluation               equal to the oracle's count from its flags
The accuracy of the scheme itself (two-level against the upsampled far
field, both in the oracle) is pinned in tests/test_oracle_far_coarse.py.
"""
import numpy as np
import pytest
import torch

from oracle import ves
from ves3d_amd import implicit as I
from ves3d_amd import workloads as W

pytestmark = pytest.mark.gpu


def cu(a):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).cuda()


def rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def _pairs(orc):
    st = orc.stokes
    p, pu, nv = st.p, st.pu, st.nv
    ng, mu = 2 * p * (p + 1), 2 * pu * (pu + 1)
    ng2 = 2 * (p // 2) * (p // 2 + 1)
    fl = st.coarse_flags
    nw = fl.shape[0]
    fine = 0
    for w in range(nw):
        a, b = (w * 64) // ng, min(w * 64 + 63, nv * ng - 1) // ng
        for g in range(nv):
            if not fl[w, g] and not (a == g and b == g):
                fine += 1
    return 64.0 * (float(np.sum(fl == 1)) * ng + fine * (mu + 2) + float(np.sum(fl == 2)) * ng2)


@pytest.mark.parametrize("nv,p,kappa,kappa2,close", [(32, 8, 3.0, 0.0, False),
                                                     (24, 6, 2.0, 0.0, False),
                                                     (16, 8, 2.5, 0.0, True),
                                                     (32, 8, 2.0, 3.0, False),
                                                     (16, 8, 2.0, 3.5, True)])
def test_two_level_far_field_vs_oracle(hip, nv, p, kappa, kappa2, close):
    """close: two vesicles moved together so that near-singular pairs occur
    (their far-field part uses the upsampled sources of the near vesicle,
    which the near-singular correction subtracts)."""
    if close:
        from test_step_forms_gpu import workload
        X = workload(nv, p, 201)
    else:
        X = W.implicit_step_case(nv, p, 0.2, seed=55)
    orc = ves.Implicit(X, p, dt=1e-3, shear=1.0, r_near=0.1, far_coarse=kappa,
                       far_coarse2=kappa2)
    fl = orc.stokes.coarse_flags
    assert 0.2 < np.mean(fl > 0) < 0.95, np.mean(fl > 0)
    assert not kappa2 or 0.1 < np.mean(fl == 2) < 0.9, np.mean(fl == 2)
    assert not close or len(orc.stokes.near.pairs) > 0
    print("multi-level far field nv=%d p=%d kappa=%g kappa2=%g: level 1 %.3f, level 2 %.3f, "
          "%d near pairs" % (nv, p, kappa, kappa2, np.mean(fl == 1), np.mean(fl == 2),
                             len(orc.stokes.near.pairs)), flush=True)
    tol = 2e-8 if len(orc.stokes.near.pairs) else 5e-9
    x = np.random.default_rng(3).standard_normal(4 * nv * p * (p + 2))
    b0, y0 = orc.rhs(), orc.matvec(x)
    for precision in (0, 2):
        st = I.Stepper(nv, p, dt=1e-3, shear_rate=1.0, precision=precision, precond=1,
                       r_near=0.1, far_coarse=kappa, far_coarse2=kappa2, post=False)
        st.set_state(cu(X))
        st.prepare()
        b = st.rhs().cpu().numpy()
        y = st.matvec(cu(x)).cpu().numpy()
        b2, y2 = [t.cpu().numpy() for t in st.rhs_matvec(cu(x))]
        if precision == 0:
            assert rel(b, b0) < tol, rel(b, b0)
            assert rel(y, y0) < tol, rel(y, y0)
            assert rel(b2, b0) < tol and rel(y2, y0) < tol
        else:
            for u, v in ((b, b0), (y, y0), (b2, b0), (y2, y0)):
                e = np.linalg.norm(u - v) / np.linalg.norm(v)
                assert e < 2e-5, e
        s = st.step()
        assert s["status"] == 100
        assert s["far_pairs"] == _pairs(orc), (s["far_pairs"], _pairs(orc))
        st.close()

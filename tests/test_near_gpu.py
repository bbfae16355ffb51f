"""GPU parity of the near-singular correction (ves3d_amd/csrc/near.hip)
against the oracle (oracle/near.py) through the C ABI, and of the implicit
step with near pairs present.

Tolerances: pair set identical; repulsion force 1e-9 relative; near velocity,
and the right-hand side / matvec with near pairs, 2e-8 relative: the patch
projection is a Newton iteration that stops at a 1e-8 tangential residual,
and the two sides reach that point along paths that differ in rounding (the
same formulas in a different association order), which moves the projected
point, and so the interpolated velocity, at the 1e-9..1e-8 level.
"""
import ctypes

import numpy as np
import pytest
import torch

from oracle import ves
from ves3d_amd import call, ptr
from ves3d_amd import implicit as I
from ves3d_amd import workloads as W

pytestmark = pytest.mark.gpu


def cu(a):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).cuda()


def rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def _close_triple(p, seed, gaps=(0.04, 0.08)):
    X = W.perturbed_shapes(3, p, amp=0.05, seed=seed)
    X[1] += np.array([2.0 + gaps[0], 0.0, 0.1])[:, None]
    X[2] += np.array([0.9, 1.9 + gaps[1], -0.2])[:, None]
    return X


@pytest.mark.parametrize("p,seed", [(6, 3), (8, 4), (12, 5)])
def test_near_pairs_repulsion_velocity(hip, p, seed):
    pu = 2 * p
    X = _close_triple(p, seed)
    nv, ng = X.shape[0], X.shape[2]
    st = ves.StokesSurface(X, p, pu, near=True, repul_dist=5e-2)
    nr = st.near
    assert len(nr.pairs) > 0
    h = ctypes.c_void_p()
    call("ves3d_hip_near_create", p, pu, nv, 5e-2, 0.0, ctypes.byref(h))
    try:
        src, trg = cu(st.scoord_far), cu(st.trg)
        print("near set_coord p=%d seed=%d pairs=%d" % (p, seed, len(nr.pairs)), flush=True)
        call("ves3d_hip_near_set_coord", h, ptr(src), ptr(trg), None)
        npairs, rn = ctypes.c_size_t(), ctypes.c_double()
        call("ves3d_hip_near_info", h, ctypes.byref(npairs), ctypes.byref(rn))
        assert npairs.value == len(nr.pairs)
        assert rn.value == pytest.approx(nr.r_near, rel=1e-13)
        f = torch.zeros(nv * 3 * ng, dtype=torch.float64, device="cuda")
        call("ves3d_hip_near_force_repul", h, ptr(f), None)
        torch.cuda.synchronize()
        assert rel(f.cpu().numpy().reshape(nv, 3, ng), st.force_repul()) < 1e-9
        dens = np.random.default_rng(seed).uniform(-1, 1, X.shape)
        s_vel, V, q, _, _, far = st.terms(dens)
        vu, vp = st.self_velocity_up(V)
        want = np.transpose(st.near_velocity(q, V).reshape(nv, ng, 3), (0, 2, 1))
        vel = torch.zeros(nv * 3 * ng, dtype=torch.float64, device="cuda")
        gq, gu, gp = cu(q), cu(vu), cu(vp)     # referenced until the kernels ran
        call("ves3d_hip_near_apply", h, ptr(gq), ptr(gu), ptr(gp), ptr(vel), None)
        torch.cuda.synchronize()
        got = vel.cpu().numpy().reshape(nv, 3, ng)
        assert rel(got, want) < 2e-8
    finally:
        call("ves3d_hip_near_destroy", h)


def test_no_pairs_when_far_apart(hip):
    p = 6
    X = W.perturbed_shapes(2, p, amp=0.05, seed=1)
    X[1] += np.array([3.5, 0.0, 0.0])[:, None]
    st = ves.StokesSurface(X, p, 2 * p, near=True)
    assert len(st.near.pairs) == 0
    h = ctypes.c_void_p()
    call("ves3d_hip_near_create", p, 2 * p, 2, 5e-2, 0.0, ctypes.byref(h))
    try:
        src, trg = cu(st.scoord_far), cu(st.trg)
        call("ves3d_hip_near_set_coord", h, ptr(src), ptr(trg), None)
        n = ctypes.c_size_t()
        call("ves3d_hip_near_info", h, ctypes.byref(n), None)
        assert n.value == 0
    finally:
        call("ves3d_hip_near_destroy", h)


@pytest.mark.parametrize("precision", [0])
def test_stepper_rhs_matvec_with_near(hip, precision):
    p = 8
    X = _close_triple(p, 9)
    dt, shear = 1e-3, 0.3
    orc = ves.Implicit(X, p, dt=dt, shear=shear)
    assert len(orc.stokes.near.pairs) > 0
    st = I.Stepper(3, p, dt=dt, shear_rate=shear, precision=precision, precond=1)
    st.set_state(cu(X))
    st.prepare()
    print("stepper rhs/matvec with near pairs=%d" % len(orc.stokes.near.pairs), flush=True)
    b = st.rhs().cpu().numpy()
    assert rel(b, orc.rhs()) < 2e-8
    x = np.random.default_rng(3).standard_normal(st.n)
    y = st.matvec(cu(x)).cpu().numpy()
    assert rel(y, orc.matvec(x)) < 2e-8
    st.close()

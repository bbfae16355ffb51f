"""Per-vesicle properties and gravity in the step (VesicleProperties,
src/VesicleProps.cc:19-55; InterfacialForce::gravityForce,
src/InterfacialForce.cc:126-184) against the oracle, and one step of the
reference's own polydisperse sedimentation suspension
(experiment/polydisperse_config.in: precomputed/geometry_spec_polysed_512.txt,
vesprops_rand_polysed_512.txt, the p = 16 shape gallery's first shape, no
background flow, dt 1e-2, tolerance 1e-5).

Bars: FP64 right-hand side and matvec 1e-9 relative (2e-8 with near pairs);
step velocity 1e-5 against the oracle's solve to 1e-10.  The 512-vesicle
suspension has no oracle at that size: both forms converge, both meet the
tolerance under the FP64 operator, their velocities agree to 1e-4, and the
suspension sediments (mean vertical velocity below zero, the large vesicles
faster than the small ones).
"""
import os

import numpy as np
import pytest
import torch

from oracle import ves
from ves3d_amd import Ves3dError
from ves3d_amd import implicit as I
from ves3d_amd import workloads as W

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def cu(a):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).cuda()


def rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def test_props_and_gravity_vs_oracle(hip):
    nv, p = 4, 8
    X = W.implicit_step_case(nv, p, 0.2, seed=61)
    kb = np.array([0.01, 0.04, 0.08, 0.02])
    rho = np.array([0.0, 1.0, 0.5, 2.0])
    g = np.array([0.1, 0.0, -1.0])
    orc = ves.Implicit(X, p, dt=1e-2, kb=kb, shear=0.0, r_near=0.1, excess_density=rho,
                       gravity=g)
    tol = 2e-8 if len(orc.stokes.near.pairs) else 1e-9
    st = I.Stepper(nv, p, dt=1e-2, shear_rate=0.0, precision=0, precond=1, tol=1e-6,
                   max_iter=300, r_near=0.1, post=False)
    st.set_vesicle_props(bending_modulus=kb, excess_density=rho, gravity=g)
    st.set_state(cu(X))
    st.prepare()
    b = st.rhs().cpu().numpy()
    b0 = orc.rhs()
    assert rel(b, b0) < tol, rel(b, b0)
    x = np.random.default_rng(4).standard_normal(st.n)
    y = st.matvec(cu(x)).cpu().numpy()
    assert rel(y, orc.matvec(x)) < tol
    x0, _, it0, s0 = ves.bicgstab(orc.matvec, b0, np.zeros_like(b0), 1e-10, 500, orc.precond)
    assert s0 == "success"
    u0, _ = orc.unpack(x0)
    s = st.step()
    U = st.get_state()[1].cpu().numpy()
    e = np.linalg.norm(U - u0) / np.linalg.norm(u0)
    print("props + gravity: %d iterations (oracle %d), velocity err %.2e" %
          (s["iterations"], it0, e), flush=True)
    assert s["status"] == 100 and e < 1e-5
    # gravity alone moves the heavy vesicles down
    assert np.mean(U[3, 2]) < np.mean(U[0, 2])
    with pytest.raises(Ves3dError) as ei:
        st.set_vesicle_props(viscosity_contrast=[1.0, -2.0, 1.0, 1.0])
    assert ei.value.code == 10
    st.close()


def test_polydisperse_sedimentation_step(hip):
    p, nv = 16, 512
    X = W.suspension_from_files(os.path.join(GOLD, "shape_gallery_16_shape0.txt"),
                                os.path.join(GOLD, "geometry_spec_polysed_512.txt"), p)
    kb, vc, rho = W.read_vesicle_props(os.path.join(GOLD, "vesprops_rand_polysed_512.txt"), nv)
    scale = np.loadtxt(os.path.join(GOLD, "geometry_spec_polysed_512.txt"))[:, 1]
    common = dict(dt=1e-2, shear_rate=0.0, tol=1e-5, max_iter=300, post=False)
    U = {}
    for name, kw in (("plain", dict(precision=0, precond=1)),
                     ("fast", dict(precision=2, precond=1, inner_max_iter=6, inner_tol=1e-2,
                                   far_coarse=3.0))):
        st = I.Stepper(nv, p, **kw, **common)
        st.set_vesicle_props(bending_modulus=kb, viscosity_contrast=vc, excess_density=rho)
        st.set_state(cu(X))
        s = st.step()
        _, u, sig = st.get_state()
        st.close()
        chk = I.Stepper(nv, p, precision=0, precond=1, **common)
        chk.set_vesicle_props(bending_modulus=kb, viscosity_contrast=vc, excess_density=rho)
        chk.set_state(cu(X), u, sig)
        r64 = chk.residual()
        chk.close()
        print("polydisperse sedimentation, %s: %d iterations, %.0f ms, %d near pairs, FP64 relres "
              "%.2e" % (name, s["iterations"], s["ms_total"], s["near_pairs"], r64), flush=True)
        assert s["status"] == 100 and r64 <= 1e-5 * 1.05, (name, s, r64)
        U[name] = u.cpu().numpy()
    e = np.linalg.norm(U["fast"] - U["plain"]) / np.linalg.norm(U["plain"])
    assert e < 1e-4, e
    uz = U["plain"][:, 2].mean(axis=1)
    big, small = uz[scale > 1.0].mean(), uz[scale < 1.0].mean()
    print("mean vertical velocity: large %.4f, small %.4f" % (big, small), flush=True)
    assert big < small < 0.0

"""GPU parity of the SphericalHarmonics transforms, rotations, singular
self-interaction build and self apply against the numpy FP64 oracle.

Tolerances (max-norm error relative to the max of the oracle output):
  FP64 transforms / rotations     1e-12
  FP64 singular operators         1e-9  (the host-built MatRotate differs from the
                                          oracle's by O(1e-8) in entries fed by the
                                          rotated-pole grid point, see test_oracle_sph)
  FP32 transforms                 1e-5
  FP32 singular operators         1e-5 (vs FP64 oracle)
"""
import numpy as np
import pytest
import torch

from oracle import sph as O
from oracle import ves
from ves3d_amd import sph as G
from ves3d_amd import workloads as W

pytestmark = pytest.mark.gpu
RNG = np.random.default_rng(21)


def cu(a, dt=torch.float64):
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda", dt)


def rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


@pytest.mark.parametrize("p0,p1,N", [(6, 6, 5), (8, 16, 3), (12, 12, 7), (16, 32, 11), (16, 16, 1)])
@pytest.mark.parametrize("dt,tol", [(torch.float64, 1e-12), (torch.float32, 1e-5)])
def test_shc2grid(hip, p0, p1, N, dt, tol):
    S = RNG.standard_normal((N, O.ncoef(p0) if hasattr(O, "ncoef") else p0 * (p0 + 2)))
    X, Xt, Xp = O.shc2grid(S, p0, p1, True, True)
    g = lambda: torch.zeros((N, 2 * p1 * (p1 + 1)), dtype=dt, device="cuda")
    gx, gt, gp = g(), g(), g()
    G.shc2grid(cu(S, dt), N, p0, p1, gx, gt, gp)
    assert rel(gx.double().cpu().numpy(), X) < tol
    assert rel(gt.double().cpu().numpy(), Xt) < tol
    assert rel(gp.double().cpu().numpy(), Xp) < tol
    gx2 = g()
    G.shc2grid(cu(S, dt), N, p0, p1, gx2)  # no derivative outputs
    assert rel(gx2.double().cpu().numpy(), X) < tol


@pytest.mark.parametrize("p0,p1,N", [(6, 6, 5), (16, 8, 3), (12, 12, 7), (32, 16, 13), (24, 24, 2)])
@pytest.mark.parametrize("dt,tol", [(torch.float64, 1e-12), (torch.float32, 1e-5)])
def test_grid2shc_and_transpose(hip, p0, p1, N, dt, tol):
    X = RNG.standard_normal((N, 2 * p0 * (p0 + 1)))
    S = O.grid2shc(X, p0, p1)
    T = O.shc2grid_transpose(X, p0, p1)
    gs = torch.zeros((N, p1 * (p1 + 2)), dtype=dt, device="cuda")
    G.grid2shc(cu(X, dt), N, p0, p1, gs)
    assert rel(gs.double().cpu().numpy(), S) < tol
    G.shc2grid_transpose(cu(X, dt), N, p0, p1, gs)
    assert rel(gs.double().cpu().numpy(), T) < tol


@pytest.mark.parametrize("p,N", [(6, 4), (16, 9)])
def test_shc2pole(hip, p, N):
    S = RNG.standard_normal((N, p * (p + 2)))
    P = O.shc2pole(S, p)
    gp = torch.zeros((N, 2), dtype=torch.float64, device="cuda")
    G.shc2pole(cu(S), N, p, gp)
    assert rel(gp.cpu().numpy(), P) < 1e-13


@pytest.mark.parametrize("p,N,dof", [(6, 2, 3), (8, 1, 12), (16, 2, 3), (32, 1, 1)])
def test_rotations(hip, p, N, dof):
    """RotateAll / RotateTranspose against the oracle; p = 32 in FP64 takes
    the scatter kernel's unstaged path (four rows exceed its LDS staging)."""
    nc = p * (p + 2)
    S = RNG.standard_normal((N * dof, nc))
    R = O.rotate_all(S, p, dof)
    out = torch.zeros(R.size, dtype=torch.float64, device="cuda")
    G.rotate_all(cu(S), N, p, dof, out)
    # rotation matrices are rebuilt on the host; entries fed by the rotated-pole
    # grid point carry O(1e-8) rounding noise in both (see test_oracle_sph)
    assert rel(out.cpu().numpy().reshape(R.shape), R) < 1e-7
    Y = RNG.standard_normal(R.shape)
    T = O.rotate_transpose(Y, p, dof)
    out2 = torch.zeros(R.size, dtype=torch.float64, device="cuda")
    G.rotate_transpose(cu(Y), N, p, dof, out2)
    assert rel(out2.cpu().numpy().reshape(T.shape), T) < 1e-7


def _singular_case(nv, p0, seed):
    return W.perturbed_shapes(nv, p0, amp=0.1, seed=seed)


@pytest.mark.parametrize("nv,p0,p1", [(3, 6, 12), (2, 8, 16), (1, 12, 24)])
def test_singular_build_f64(hip, nv, p0, p1):
    X = _singular_case(nv, p0, seed=nv + p0)
    SL, DL = O.stokes_singular_integ(X, p0, p1)
    n = 3 * p0 * (p0 + 2)
    gsl = torch.zeros((nv, n, n), dtype=torch.float64, device="cuda")
    gdl = torch.zeros((nv, n, n), dtype=torch.float64, device="cuda")
    G.singular_build(cu(X), nv, p0, p1, gsl, gdl)
    e1 = rel(gsl.cpu().numpy(), SL)
    e2 = rel(gdl.cpu().numpy(), DL)
    assert e1 < 1e-9 and e2 < 1e-9, (e1, e2)
    # SL only
    g2 = torch.zeros_like(gsl)
    G.singular_build(cu(X), nv, p0, p1, g2, None)
    assert rel(g2.cpu().numpy(), SL) < 1e-9


def test_singular_build_f32(hip):
    nv, p0, p1 = 2, 8, 16
    X = _singular_case(nv, p0, seed=5)
    SL, DL = O.stokes_singular_integ(X, p0, p1)
    n = 3 * p0 * (p0 + 2)
    gsl = torch.zeros((nv, n, n), dtype=torch.float32, device="cuda")
    gdl = torch.zeros((nv, n, n), dtype=torch.float32, device="cuda")
    G.singular_build(cu(X), nv, p0, p1, gsl, gdl)
    assert rel(gsl.double().cpu().numpy(), SL) < 1e-5
    assert rel(gdl.double().cpu().numpy(), DL) < 1e-5


@pytest.mark.parametrize("p0", [12, 16, 24])
def test_singular_closed_forms_gpu(hip, p0):
    """Size-independent property at the workload's order: on unit spheres
    SL[c] = (2/3) c, DL[c] = -c/2 (the grid's outward normal, the
    reference orientation), SL[n] = 0 (p1 = 2 p0)."""
    nv = 2
    Xs, _ = W.sphere_grid(p0)
    X = np.stack([Xs, Xs * 1.0])
    n = 3 * p0 * (p0 + 2)
    nc = p0 * (p0 + 2)
    gsl = torch.zeros((nv, n, n), dtype=torch.float64, device="cuda")
    gdl = torch.zeros((nv, n, n), dtype=torch.float64, device="cuda")
    G.singular_build(cu(X), nv, p0, 2 * p0, gsl, gdl)
    c = np.array([0.3, -1.1, 0.7])
    f = np.repeat(c, Xs.shape[1]).reshape(3, -1)
    F = O.grid2shc(f, p0, p0).reshape(-1)
    Fn = O.grid2shc(Xs, p0, p0).reshape(-1)
    for v in range(nv):
        SLv = gsl[v].cpu().numpy()
        DLv = gdl[v].cpu().numpy()
        V = O.shc2grid((F @ SLv).reshape(3, nc), p0, p0)
        Vd = O.shc2grid((F @ DLv).reshape(3, nc), p0, p0)
        Vn = O.shc2grid((Fn @ SLv).reshape(3, nc), p0, p0)
        assert np.max(np.abs(V - 2.0 / 3.0 * f)) < 1e-9
        assert np.max(np.abs(Vd + 0.5 * f)) < 1e-9
        assert np.max(np.abs(Vn)) < 1e-9


@pytest.mark.parametrize("p0,sl_tol", [(16, 5e-8), (24, 1e-8)])
def test_singular_identities_on_ellipsoids_gpu(hip, p0, sl_tol):
    """Shape-sensitive identities at the workloads' orders, on two rotated
    ellipsoids (semi-axes 1, 1, 0.6 +- 0.05), where a target rotated to the
    wrong pole would break them: DL[c] = -c/2 on the surface and SL[n] = 0
    (n truncated at p0, whose error decays with p0), p1 = 2 p0, for the
    FP64 build and, to the FP32 build's accuracy, the FP32 one."""
    nv = 2
    X = W.implicit_step_case(nv, p0, 0.2, seed=9)
    X = X - X.mean(axis=2, keepdims=True)
    n = 3 * p0 * (p0 + 2)
    nc = p0 * (p0 + 2)
    c = np.array([0.3, -1.1, 0.7])
    f = np.repeat(c, X.shape[2]).reshape(3, -1)
    F = O.grid2shc(f, p0, p0).reshape(-1)
    nrm = ves.Surface(X, p0).normal
    for dt, dl_tol, sl_bar in ((torch.float64, 1e-9, sl_tol), (torch.float32, 1e-4, 1e-4)):
        print("singular identities p0=%d %s" % (p0, dt), flush=True)
        gsl = torch.zeros((nv, n, n), dtype=dt, device="cuda")
        gdl = torch.zeros((nv, n, n), dtype=dt, device="cuda")
        G.singular_build(cu(X), nv, p0, 2 * p0, gsl, gdl)
        for v in range(nv):
            SLv = gsl[v].double().cpu().numpy()
            DLv = gdl[v].double().cpu().numpy()
            Vd = O.shc2grid((F @ DLv).reshape(3, nc), p0, p0)
            Fn = O.grid2shc(nrm[v], p0, p0).reshape(-1)
            Vn = O.shc2grid((Fn @ SLv).reshape(3, nc), p0, p0)
            assert np.max(np.abs(Vd + 0.5 * f)) < dl_tol, (dt, v)
            assert np.max(np.abs(Vn)) < sl_bar, (dt, v)
        del gsl, gdl


@pytest.mark.parametrize("dt,tol", [(torch.float64, 1e-13), (torch.float32, 1e-6)])
def test_self_apply(hip, dt, tol):
    nv, p = 5, 8
    n = 3 * p * (p + 2)
    SL = RNG.standard_normal((nv, n, n))
    DL = RNG.standard_normal((nv, n, n))
    F = RNG.standard_normal((nv, n))
    Gd = RNG.standard_normal((nv, n))
    ref = np.einsum("vr,vrc->vc", F, SL) + np.einsum("vr,vrc->vc", Gd, DL)
    V = torch.zeros((nv, n), dtype=torch.float64, device="cuda")
    G.self_apply(cu(SL, dt), cu(F), cu(DL, dt), cu(Gd), nv, p, V)
    assert rel(V.cpu().numpy(), ref) < tol
    ref1 = np.einsum("vr,vrc->vc", F, SL)
    G.self_apply(cu(SL, dt), cu(F), None, None, nv, p, V)
    assert rel(V.cpu().numpy(), ref1) < tol


@pytest.mark.parametrize("p0", [16, 8])
def test_singular_build_fused_pole_stage(hip, p0):
    """The FP64 build's fused pole kernel + transposed synthesis (the default
    where p1 = 2 p0, p0 <= 16) against the separate kernel and grouped GEMMs,
    SL and DL.  Asserts that the fused stage applies on this device."""
    import ctypes
    from ves3d_amd import call
    nv, p1, dt, tol = 3, 2 * p0, torch.float64, 1e-12
    rows, threads, groups = ctypes.c_int(0), ctypes.c_int(0), ctypes.c_int(0)
    call("ves3d_hip_pole_kernel_shape", p0, p1, ctypes.byref(rows), ctypes.byref(threads),
         ctypes.byref(groups))
    print("fused pole stage p0=%d: %d rows per pass, %d threads, %d workgroups"
          % (p0, rows.value, threads.value, groups.value), flush=True)
    assert rows.value >= 1 and threads.value >= 64 and groups.value >= 1
    X = _singular_case(nv, p0, seed=41)
    n = 3 * p0 * (p0 + 2)
    out = {}
    try:
        for fused in (1, 0):
            call("ves3d_hip_set_pole_kernel", fused)
            sl = torch.zeros((nv, n, n), dtype=dt, device="cuda")
            dl = torch.zeros((nv, n, n), dtype=dt, device="cuda")
            G.singular_build(cu(X), nv, p0, p1, sl, dl)
            torch.cuda.synchronize()
            out[fused] = (sl.double().cpu().numpy(), dl.double().cpu().numpy())
    finally:
        call("ves3d_hip_set_pole_kernel", 1)
    e_sl, e_dl = rel(out[1][0], out[0][0]), rel(out[1][1], out[0][1])
    print("fused vs separate pole stage (p0=%d): SL %.2e, DL %.2e" % (p0, e_sl, e_dl), flush=True)
    assert e_sl < tol and e_dl < tol


@pytest.mark.parametrize("p0,fused,dt", [(16, 1, torch.float64), (16, 0, torch.float64),
                                         (12, 1, torch.float64), (16, 1, torch.float32)])
def test_singular_operator_block_symmetry(hip, p0, fused, dt):
    """The built SL and DL operators' 3 x 3 blocks of Ncoef x Ncoef satisfy
    B_ab = B_ba exactly (block (a, b) holds the rows of the symmetric tensor
    component c(a, b)), which the step's self apply relies on to read only
    the blocks a <= b.  Fused FP64 build, separate FP64 build, and the FP32
    build (p0 = 12 takes the separate path whatever the setting)."""
    from ves3d_amd import call
    nv = 2
    nc = p0 * (p0 + 2)
    X = _singular_case(nv, p0, seed=43)
    try:
        call("ves3d_hip_set_pole_kernel", fused)
        sl = torch.zeros((nv, 3 * nc, 3 * nc), dtype=dt, device="cuda")
        dl = torch.zeros((nv, 3 * nc, 3 * nc), dtype=dt, device="cuda")
        G.singular_build(cu(X), nv, p0, 2 * p0, sl, dl)
        torch.cuda.synchronize()
    finally:
        call("ves3d_hip_set_pole_kernel", 1)
    for M in (sl.cpu().numpy(), dl.cpu().numpy()):
        assert np.max(np.abs(M)) > 0
        for a in range(3):
            for b in range(a + 1, 3):
                B_ab = M[:, a * nc:(a + 1) * nc, b * nc:(b + 1) * nc]
                B_ba = M[:, b * nc:(b + 1) * nc, a * nc:(a + 1) * nc]
                assert np.array_equal(B_ab, B_ba), (a, b)

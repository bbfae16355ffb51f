"""GPU parity of the implicit step against the numpy FP64 oracle
(oracle/ves.py): right-hand side, Galerkin matvec, diagonal spectral
preconditioner, the device BiCGStab, and one full step.

Tolerances (relative, max-norm unless stated):
  FP64 rhs / matvec / precond   1e-9
  FP64 step solution            1e-5 (both solves stop at relres 1e-6)
  mixed-precision matvec        2e-5 against the FP64 oracle (precision 1:
                                FP32 singular operator + FP32c far field;
                                2: FP64 singular operator + FP32c far field)
"""
import ctypes

import numpy as np
import pytest
import torch

from oracle import ves
from ves3d_amd import implicit as I
from ves3d_amd import workloads as W

pytestmark = pytest.mark.gpu


def cu(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def _case(nv, p, seed, volfrac=0.05):
    shapes = W.perturbed_shapes(nv, p, amp=0.05, seed=seed)
    return W.suspension(shapes, volfrac, seed=seed)


@pytest.fixture(params=[1, 0, 2], ids=["far-scalar", "far-lds", "far-scalar2"])
def far_kernel(request, hip):
    """The kernels of the far field between vesicles (scalar-cache source
    records; LDS-staged source tiles; scalar-cache records with two targets
    per lane in the FP32 form)."""
    from ves3d_amd import call
    call("ves3d_hip_set_far_kernel", request.param)
    yield request.param
    call("ves3d_hip_set_far_kernel", 1)


@pytest.mark.parametrize("nv,p", [(2, 6), (3, 8)])
def test_rhs_matvec_precond_f64(hip, far_kernel, nv, p):
    X = _case(nv, p, seed=10 + nv)
    dt, shear = 1e-3, 0.3
    orc = ves.Implicit(X, p, dt=dt, shear=shear)
    st = I.Stepper(nv, p, dt=dt, shear_rate=shear, precision=0, precond=1)
    st.set_state(cu(X))
    st.prepare()
    b = st.rhs().cpu().numpy()
    b0 = orc.rhs()
    # 2e-8 where the case has near-singular pairs (see test_near_gpu.py)
    tol = 2e-8 if len(orc.stokes.near.pairs) else 1e-9
    assert rel(b, b0) < tol, rel(b, b0)
    x = np.random.default_rng(nv).standard_normal(st.n)
    y = st.matvec(cu(x)).cpu().numpy()
    y0 = orc.matvec(x)
    assert rel(y, y0) < tol, rel(y, y0)
    z = st.precond(cu(x)).cpu().numpy()
    assert rel(z, orc.precond(x)) < 1e-14
    st.close()


@pytest.mark.parametrize("precision", [1, 2])
def test_matvec_mixed_precision(hip, far_kernel, precision):
    nv, p = 3, 8
    X = _case(nv, p, seed=7)
    orc = ves.Implicit(X, p, dt=1e-3, shear=0.3)
    st = I.Stepper(nv, p, dt=1e-3, shear_rate=0.3, precision=precision, precond=1)
    st.set_state(cu(X))
    st.prepare()
    x = np.random.default_rng(1).standard_normal(st.n)
    y = st.matvec(cu(x)).cpu().numpy()
    e = np.linalg.norm(y - orc.matvec(x)) / np.linalg.norm(orc.matvec(x))
    assert e < 2e-5, e
    st.close()


def test_device_bicgstab_callbacks(hip):
    """ves3d_hip_bicgstab with C callbacks: solve D x = b for the diagonal
    spectral scaling D, compared with b / D."""
    nv, p = 2, 6
    X = _case(nv, p, seed=3)
    st = I.Stepper(nv, p, precision=0, precond=1)
    st.set_state(cu(X))
    lib = st  # keep alive
    fn = __import__("ves3d_amd").load().ves3d_hip_stepper_precond

    def apply_d(ctx, x, y):
        return fn(lib.h, x, y)

    b = torch.from_numpy(np.random.default_rng(2).standard_normal(st.n)).cuda()
    x = torch.zeros_like(b)
    it, res, status = I.bicgstab(apply_d, x, b, 1e-10, 200)
    assert status == 100 and res < 1e-10
    d = st.precond(torch.ones_like(b)).cpu().numpy()
    assert rel(x.cpu().numpy(), b.cpu().numpy() / d) < 1e-8
    st.close()


@pytest.mark.parametrize("restart,flexible", [(8, False), (8, True), (40, True)])
def test_device_fgmres_dense(hip, restart, flexible):
    """ves3d_hip_fgmres with C callbacks on a dense nonsymmetric system
    (n = 200, eigenvalues within about 2 of 4), the operator applied by the
    library's GEMM, against numpy.linalg.solve.  flexible: the
    preconditioner alternates between two different diagonal scalings, so it
    changes from one application to the next (what the flexible form
    allows); restart 8 forces several cycles.  The returned residual is the
    explicit one, compared with ||b - A x|| / ||b|| recomputed on the host."""
    from ves3d_amd import load
    lib = load()
    rng = np.random.default_rng(11)
    n = 200
    A = np.diag(4.0 + rng.uniform(-1, 1, n)) + rng.standard_normal((n, n)) * (1.2 / np.sqrt(n))
    b = rng.standard_normal(n)
    d = np.diag(A)
    P1 = np.diag(1.0 / d)
    P2 = np.diag(1.0 / (d * rng.uniform(0.5, 2.0, n)))
    gA, gP1, gP2 = (cu(np.ascontiguousarray(M.T)) for M in (A, P1, P2))  # column-major
    calls = [0]

    def gemv(M, x, y):
        return lib.ves3d_hip_gemm_f64(b"N", b"N", n, 1, n, 1.0, M.data_ptr(), n, x, n, 0.0, y, n)

    def apply_a(ctx, x, y):
        return gemv(gA, x, y)

    def apply_p(ctx, x, y):
        calls[0] += 1
        return gemv(gP2 if flexible and calls[0] % 2 == 0 else gP1, x, y)

    x = torch.zeros(n, dtype=torch.float64, device="cuda")
    it, res, status = I.fgmres(apply_a, x, cu(b), 1e-12, 400, restart=restart, apply_p=apply_p)
    xs = x.cpu().numpy()
    want = np.linalg.solve(A, b)
    true = float(np.linalg.norm(b - A @ xs) / np.linalg.norm(b))
    print("fgmres restart %d flexible %d: %d iterations, %d preconditioner calls, relres %.2e, "
          "host relres %.2e, error %.2e" % (restart, flexible, it, calls[0], res, true,
                                            np.linalg.norm(xs - want) / np.linalg.norm(want)),
          flush=True)
    assert status == 100 and res <= 1e-12
    assert calls[0] == it
    assert true <= 1e-11 and abs(true - res) <= 1e-12
    assert np.linalg.norm(xs - want) / np.linalg.norm(want) < 1e-10
    if restart >= 40:
        assert it < 40  # no restart needed: GMRES within one cycle


def test_device_fgmres_limit_and_nan(hip):
    """The iteration limit gives status 101 with the explicit residual of the
    returned iterate; a NaN from the operator gives status 104."""
    from ves3d_amd import load
    lib = load()
    rng = np.random.default_rng(12)
    n = 64
    A = np.eye(n) * 3.0 + rng.standard_normal((n, n)) * 0.6
    b = rng.standard_normal(n)
    gA = cu(np.ascontiguousarray(A.T))

    def apply_a(ctx, x, y):
        return lib.ves3d_hip_gemm_f64(b"N", b"N", n, 1, n, 1.0, gA.data_ptr(), n, x, n, 0.0, y, n)

    x = torch.zeros(n, dtype=torch.float64, device="cuda")
    it, res, status = I.fgmres(apply_a, x, cu(b), 1e-14, 5, restart=3)
    xs = x.cpu().numpy()
    true = float(np.linalg.norm(b - A @ xs) / np.linalg.norm(b))
    assert status == 101 and it == 5 and abs(true - res) <= 1e-12 * max(1.0, true), (res, true)
    gN = cu(np.full((n, n), np.nan))

    def apply_nan(ctx, x, y):
        return lib.ves3d_hip_gemm_f64(b"N", b"N", n, 1, n, 1.0, gN.data_ptr(), n, x, n, 0.0, y, n)

    x = torch.zeros(n, dtype=torch.float64, device="cuda")
    it, res, status = I.fgmres(apply_nan, x, cu(b), 1e-10, 20, restart=10)
    assert status == 104


@pytest.mark.parametrize("precond", [0, 1])
def test_full_step_f64(hip, precond):
    nv, p = 2, 6
    X = _case(nv, p, seed=4)
    dt, shear = 1e-3, 1.0
    orc = ves.Implicit(X, p, dt=dt, shear=shear)
    b0 = orc.rhs()
    x0, res0, it0, st0 = ves.bicgstab(orc.matvec, b0, np.zeros_like(b0), 1e-6, 300,
                                      orc.precond if precond else None)
    assert st0 == "success"
    u0, s0 = orc.unpack(x0)
    st = I.Stepper(nv, p, dt=dt, shear_rate=shear, precision=0, precond=precond, tol=1e-6,
                   max_iter=300, post=False)
    st.set_state(cu(X))
    stats = st.step()
    assert stats["status"] == 100 and stats["true_relres"] < 1e-6 * 1.01, stats
    Xn, Ug, sig = st.get_state()
    U = Ug.cpu().numpy()
    assert np.linalg.norm(U - u0) / np.linalg.norm(u0) < 1e-5
    assert np.linalg.norm(Xn.cpu().numpy() - (X + dt * u0)) / np.linalg.norm(X) < 1e-9
    st.close()
    # the residual entry point reproduces the step's true-residual check
    chk = I.Stepper(nv, p, dt=dt, shear_rate=shear, precision=0, precond=precond, post=False)
    chk.set_state(cu(X), Ug, sig)
    r = chk.residual()
    assert abs(r - stats["true_relres"]) < 1e-6 * stats["true_relres"] + 1e-15, (r, stats)
    chk.close()


def test_reparam_and_area_volume_vs_oracle(hip):
    """Reparametrization and area / volume correction on the device against
    the oracle (oracle/ves.py reparam, area_volume_correction): same
    iteration counts, positions to 1e-9 relative."""
    nv, p = 3, 8
    X = _case(nv, p, seed=21, volfrac=0.02)
    st = I.Stepper(nv, p, precision=0, precond=1)
    st.set_state(cu(X))
    it = st.reparam()
    Xr, it0 = ves.reparam(X, p)
    Xg, _, _ = st.get_state()
    assert it == it0
    assert rel(Xg.cpu().numpy(), Xr) < 1e-9
    # perturb the recorded targets: the correction restores them
    a0, v0 = ves.area_volume(Xr, p, 2 * p)
    st.set_area_volume(a0 * 1.001, v0 * 0.999)
    it = st.area_volume()
    Xa, it1 = ves.area_volume_correction(Xr, p, 2 * p, a0 * 1.001, v0 * 0.999)
    Xg, _, _ = st.get_state()
    assert it == it1 and it > 0
    assert rel(Xg.cpu().numpy(), Xa) < 1e-9
    a1, v1 = ves.area_volume(Xg.cpu().numpy(), p, 2 * p)
    assert np.max(np.abs(a1 - a0 * 1.001) / a0) < 1e-11
    assert np.max(np.abs(v1 - v0 * 0.999) / v0) < 1e-11
    st.close()


def test_full_step_with_post(hip):
    nv, p = 2, 6
    X = _case(nv, p, seed=4)
    st = I.Stepper(nv, p, dt=1e-3, shear_rate=1.0, precision=0, precond=1, tol=1e-6, max_iter=300)
    st.set_state(cu(X))
    stats = st.step()
    assert stats["status"] == 100 and stats["reparam_iters"] >= 1, stats
    Xn, U, _ = st.get_state()
    Xr, _ = ves.reparam(X + 1e-3 * U.cpu().numpy(), p)
    a0, v0 = ves.area_volume(X, p, 2 * p)
    Xa, _ = ves.area_volume_correction(Xr, p, 2 * p, a0, v0)
    assert rel(Xn.cpu().numpy(), Xa) < 1e-9
    st.close()


def test_reference_cpu_case_p12(hip):
    """The plain method's CPU-runnable case (2 perturbed spheres at p = 12,
    centres (+-1.25, 0, 0), kb = 0.01, shear 0.1, dt = 1e-3, BiCGStab to
    1e-8) against the oracle's solve of the same step: velocity to 1e-6
    relative; tension to 1e-6 without each vesicle's mean and 1e-5 with it.
    Both solves stop at relres 1e-8 from a zero guess with the diagonal
    preconditioner, so they share the tension gauge; the uniform tension of
    each vesicle is weakly determined (the oracle's own 1e-8 and 1e-12 solves
    differ by 1.3e-6 in tension, almost all of it a per-vesicle constant,
    and by 2e-8 in velocity)."""
    p = 12
    X = W.reference_cpu_case(p)
    orc = ves.Implicit(X, p, dt=1e-3, kb=1e-2, shear=0.1)
    b0 = orc.rhs()
    x0, res0, it0, st0 = ves.bicgstab(orc.matvec, b0, np.zeros_like(b0), 1e-8, 300, orc.precond)
    assert st0 == "success"
    u0, s0 = orc.unpack(x0)
    st = I.Stepper(2, p, dt=1e-3, shear_rate=0.1, bending_modulus=1e-2, precision=0, precond=1,
                   tol=1e-8, max_iter=300, post=False)
    st.set_state(cu(X))
    print("reference CPU case p=12: oracle %d iterations" % it0, flush=True)
    stats = st.step()
    assert stats["status"] == 100 and stats["true_relres"] < 1e-8 * 1.05, stats
    Xn, U, sig = st.get_state()
    U, sig = U.cpu().numpy(), sig.cpu().numpy()
    assert np.linalg.norm(U - u0) / np.linalg.norm(u0) < 1e-6
    assert np.linalg.norm(sig - s0) / np.linalg.norm(s0) < 1e-5
    d = (sig - sig.mean(axis=1, keepdims=True)) - (s0 - s0.mean(axis=1, keepdims=True))
    assert np.linalg.norm(d) / np.linalg.norm(s0) < 1e-6
    assert np.max(np.abs(Xn.cpu().numpy() - (X + 1e-3 * U))) < 1e-14
    st.close()


@pytest.mark.parametrize("precision,tol", [(0, 1e-12), (1, 1e-6)])
def test_rhs_matvec_one_pass(hip, far_kernel, precision, tol):
    """The right-hand side and a matvec from one two-density far-field pass
    equal the separate evaluations (to rounding; the FP32 form to its own
    accuracy), with near pairs present."""
    from test_near_gpu import _close_triple
    p = 8
    X = _close_triple(p, 9)
    st = I.Stepper(3, p, dt=1e-3, shear_rate=0.3, precision=precision, precond=1)
    st.set_state(cu(X))
    st.prepare()
    x = cu(np.random.default_rng(5).standard_normal(st.n))
    b1, y1 = st.rhs_matvec(x)
    b0 = st.rhs()
    y0 = st.matvec(x)
    assert rel(b1.cpu().numpy(), b0.cpu().numpy()) < tol
    assert rel(y1.cpu().numpy(), y0.cpu().numpy()) < tol
    st.close()


def test_extrapolated_initial_guess(hip):
    """extrapolate = 1 starts the solve from 2 x_{n-1} - x_{n-2}: the same
    trajectory to the solver tolerance, and no more iterations once two
    solutions exist."""
    nv, p = 4, 8
    X = W.implicit_step_case(nv, p, 0.15, seed=12)
    runs = {}
    for ex in (False, True):
        st = I.Stepper(nv, p, dt=1e-3, shear_rate=1.0, precision=0, precond=1, tol=1e-8,
                       max_iter=300, post=False, extrapolate=ex)
        st.set_state(cu(X))
        its = []
        for k in range(4):
            print("extrapolate %d step %d" % (ex, k), flush=True)
            s = st.step()
            assert s["status"] == 100 and s["true_relres"] < 1e-8 * 1.05, s
            its.append(s["iterations"])
        runs[ex] = (its, st.get_state()[1].cpu().numpy())
        st.close()
    (i0, U0), (i1, U1) = runs[False], runs[True]
    assert i1[:2] == i0[:2]
    assert all(a <= b for a, b in zip(i1[2:], i0[2:])), (i0, i1)
    assert np.linalg.norm(U1 - U0) / np.linalg.norm(U0) < 1e-6

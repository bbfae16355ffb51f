"""The two benchmarked forms of the implicit step against the oracle, and
the step's failure handling.

Workload: perturbed ellipsoids (semi-axes 1, 1, 0.6 +- 0.05, random
orientations) at 20% volume fraction, shear rate 1.0, dt = 1e-3, near zone
r_near = 0.1, BiCGStab to 1e-6 (the full-step workload), with two vesicles
moved so that grid points of each lie 0.05 from the other's surface
(near-singular pairs occur).

Forms:
  plain   : FP64 everywhere, diagonal spectral preconditioner;
  learned : precision 2 (FP32-compensated far field, FP64 singular
            operator), the learned preconditioner (trained in the test on
            pairs from plain solves of other suspensions) inside inner
            BiCGStab solves (6 iterations, 1e-2), as in the benchmark.

Each form runs with either outer solver (BiCGStab with its true-residual
check, or flexible GMRES with its explicitly formed residual).

Bars, every form, against the oracle's solve of the same system to 1e-10
(oracle/ves.py Implicit + bicgstab):
  velocity                        1e-5 relative (L2)
  residual under the FP64 operator 1e-6 (the step's own check, recomputed
                                  by a precision-0 stepper's residual())
  learned form                    no fallback
At 512 vesicles, p = 16 (no oracle at that size): the plain form and the
headline form with either solver converge, meet 1e-6 under the FP64
operator, and their velocities agree to 1e-5.
"""
import numpy as np
import pytest
import torch

from oracle import ves
from ves3d_amd import implicit as I
from ves3d_amd import training as T
from ves3d_amd import workloads as W

pytestmark = pytest.mark.gpu

COMMON = dict(dt=1e-3, shear_rate=1.0, bending_modulus=1e-2, tol=1e-6, max_iter=300,
              r_near=0.1, post=False)


def cu(a):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).cuda()


def _min_dist(A, B):
    d2 = np.sum((A[:, :, None] - B[:, None, :]) ** 2, axis=0)
    return float(np.sqrt(d2.min()))


def close_pair(X, p, a, b, gap):
    """Moves vesicle b along the line of centres until the closest grid
    point of either vesicle is `gap` from the other's surface (surfaces
    resampled to order 4p); the surfaces then stay at least gap / 3 apart."""
    X = np.array(X, copy=True)
    up = ves.resample_vec(X, p, 4 * p)
    ca, cb = X[a].mean(axis=1), X[b].mean(axis=1)
    d = (cb - ca) / np.linalg.norm(cb - ca)

    def dist(t):
        sh = t * d[:, None]
        return min(_min_dist(X[a], up[b] - sh), _min_dist(X[b] - sh, up[a]))

    lo, hi = 0.0, 0.0
    while dist(hi) > gap:
        lo, hi = hi, hi + 0.05
    for _ in range(50):
        mid = 0.5 * (lo + hi)
        if dist(mid) > gap:
            lo = mid
        else:
            hi = mid
    X[b] -= lo * d[:, None]
    up[b] -= lo * d[:, None]
    for c in range(X.shape[0]):
        if c != b:
            assert _min_dist(up[c], up[b]) >= gap / 3
    return X


def workload(nv, p, seed):
    X = W.implicit_step_case(nv, p, 0.20, seed=seed)
    c = X.mean(axis=2)
    dd = np.linalg.norm(c[:, None] - c[None, :], axis=2) + np.eye(nv) * 1e9
    a, b = np.unravel_index(np.argmin(dd), dd.shape)
    return close_pair(X, p, a, b, 0.05)


def fp64_relres(nv, p, X, U, sig, r_near=0.1):
    chk = I.Stepper(nv, p, precision=0, precond=1, **dict(COMMON, r_near=r_near))
    chk.set_state(cu(X), U, sig)
    r = chk.residual()
    chk.close()
    return r


def test_both_forms_vs_oracle(hip):
    nv, p = 5, 8
    X = workload(nv, p, seed=101)
    orc = ves.Implicit(X, p, dt=1e-3, kb=1e-2, shear=1.0, r_near=0.1)
    assert len(orc.stokes.near.pairs) > 0
    b0 = orc.rhs()
    x0, res0, it0, st0 = ves.bicgstab(orc.matvec, b0, np.zeros_like(b0), 1e-10, 500, orc.precond)
    assert st0 == "success"
    u0, s0 = orc.unpack(x0)
    print("oracle: %d iterations, %d near pairs" % (it0, len(orc.stokes.near.pairs)), flush=True)

    pc, info = T.train_learned_pc(8, p, [0.10, 0.175, 0.25], COMMON, width=256, capacity=16384,
                                  epochs=300, batch=256)
    print("trained on %d samples, loss %.3e" % (info["samples"], info["loss"]), flush=True)
    forms = {"plain": dict(precision=0, precond=1),
             "learned": dict(precision=2, precond=2, inner_max_iter=6, inner_tol=1e-2),
             "plain fgmres": dict(precision=0, precond=1, solver="fgmres"),
             "learned fgmres": dict(precision=2, precond=2, inner_max_iter=6, inner_tol=1e-2,
                                    solver="fgmres")}
    for name, kw in forms.items():
        st = I.Stepper(nv, p, **kw, **COMMON)
        if name.startswith("learned"):
            I.attach(st, pc, collect=False)
        st.set_state(cu(X))
        s = st.step()
        _, U, sig = st.get_state()
        st.close()
        e = float(np.linalg.norm(U.cpu().numpy() - u0) / np.linalg.norm(u0))
        r64 = fp64_relres(nv, p, X, U, sig)
        print("%s: %d iterations (%d inner), near pairs %d, true relres %.2e, FP64 relres "
              "%.2e, velocity err %.2e, fell back %d" % (name, s["iterations"], s["inner_iters"],
                                                         s["near_pairs"], s["true_relres"], r64,
                                                         e, s["fell_back"]), flush=True)
        assert s["status"] == 100 and s["near_pairs"] > 0, s
        assert s["fell_back"] == 0, s
        assert e <= 1e-5, (name, e)
        assert r64 <= 1e-6, (name, r64)
    pc.close()


def test_fallback_restarts_from_the_extrapolated_guess(hip):
    """A learned preconditioner that yields NaN makes every learned solve
    fail (BiCGStab status 104); the step then solves again from the same
    initial guess with the diagonal preconditioner.  With the extrapolated
    initial guess (third step on) that guess is 2 x_{n-1} - x_{n-2}, and the
    fallback's result must still meet the tolerance and follow the plain
    trajectory."""
    nv, p = 4, 6
    X = workload(nv, p, seed=102)
    nin = 4 * p * (p + 2)
    pc = I.LearnedPC(nin, 32, capacity=0, seed=3)
    pc.set_params(torch.full((pc.num_params,), float("nan"), dtype=torch.float32, device="cuda"))
    runs = {}
    for name, precond in (("diag", 1), ("nan-learned", 2)):
        st = I.Stepper(nv, p, precision=0, precond=precond, extrapolate=True,
                       **dict(COMMON, tol=1e-8))
        if precond == 2:
            I.attach(st, pc, collect=False)
        st.set_state(cu(X))
        for k in range(4):
            s = st.step()
            print(name, k, s["status"], s["iterations"], s["true_relres"], s["fell_back"],
                  flush=True)
            assert s["status"] == 100 and s["true_relres"] <= 1e-8 * 1.05, s
            assert s["fell_back"] == (1 if precond == 2 else 0), s
        runs[name] = st.get_state()[1].cpu().numpy()
        st.close()
    e = np.linalg.norm(runs["nan-learned"] - runs["diag"]) / np.linalg.norm(runs["diag"])
    assert e < 1e-6, e
    pc.close()


def test_unpreconditioned_fallback(hip):
    """fallback="none": a learned solve that fails (a NaN network) is
    repeated from the same initial guess by unpreconditioned BiCGStab, which
    still meets the tolerance and reaches the diagonal solve's velocity."""
    nv, p = 2, 6
    X = workload(nv, p, seed=104)
    nin = 4 * p * (p + 2)
    pc = I.LearnedPC(nin, 32, capacity=0, seed=3)
    pc.set_params(torch.full((pc.num_params,), float("nan"), dtype=torch.float32, device="cuda"))
    U = {}
    for name, precond, fb in (("diag", 1, "diagonal"), ("none", 2, "none")):
        st = I.Stepper(nv, p, precision=0, precond=precond, fallback=fb,
                       **dict(COMMON, tol=1e-6, max_iter=2000))
        if precond == 2:
            I.attach(st, pc, collect=False)
        st.set_state(cu(X))
        s = st.step()
        print("fallback %s: %d iterations, true relres %.2e, fell_back %d" %
              (name, s["iterations"], s["true_relres"], s["fell_back"]), flush=True)
        assert s["status"] == 100 and s["true_relres"] <= 1e-6 * 1.05, s
        assert s["fell_back"] == (1 if precond == 2 else 0)
        U[name] = st.get_state()[1].cpu().numpy()
        st.close()
    e = np.linalg.norm(U["none"] - U["diag"]) / np.linalg.norm(U["diag"])
    assert e < 1e-4, e
    pc.close()


def test_divergence_leaves_the_state(hip):
    """A solve that cannot converge within max_iter is a DivergenceError
    (src/InterfacialVelocity.cc:176-178): the step raises it with its
    statistics and leaves positions, velocity and tension unchanged."""
    nv, p = 3, 6
    X = workload(nv, p, seed=103)
    st = I.Stepper(nv, p, precision=0, precond=1, **dict(COMMON, tol=1e-12, max_iter=1))
    st.set_state(cu(X))
    before = [t.clone() for t in st.get_state()]
    with pytest.raises(I.StepDivergence) as ei:
        st.step()
    assert ei.value.stats["status"] == 101
    after = st.get_state()
    for a, b in zip(before, after):
        assert torch.equal(a, b)
    s = st.step(check=False)
    assert s["diverged"] and s["status"] == 101
    st.close()


def test_full_size_forms_agree(hip):
    """512 vesicles, p = 16: one step of each form from the same state, the
    learned one as the bench runs it (the stored network with BF16
    activations, inner solves, the three-level far field)."""
    nv, p = 512, 16
    X = W.implicit_step_case(nv, p)
    full = dict(COMMON, r_near=0.1)
    pc, meta = T.load_stored(p, 256, "bf16")
    assert pc is not None
    print("stored network: %d samples from %d plain-solve steps" %
          (meta["samples"], meta["plain_solve_steps"]), flush=True)
    sol = {}
    headline = dict(precision=2, precond=2, inner_max_iter=3, inner_tol=1e-1, far_coarse=3.0,
                    far_coarse2=6.0)
    for name, kw in (("plain", dict(precision=0, precond=1)),
                     ("learned", dict(headline)),
                     ("learned fgmres", dict(headline, solver="fgmres"))):
        st = I.Stepper(nv, p, **kw, **full)
        if name.startswith("learned"):
            I.attach(st, pc, collect=False)
        st.set_state(cu(X))
        s = st.step()
        _, U, sig = st.get_state()
        st.close()
        r64 = fp64_relres(nv, p, X, U, sig)
        print("%s: %d iterations, %.0f ms, true relres %.2e, FP64 relres %.2e" %
              (name, s["iterations"], s["ms_total"], s["true_relres"], r64), flush=True)
        assert s["status"] == 100 and r64 <= 1e-6, (name, s, r64)
        if name.startswith("learned"):
            assert s["fell_back"] == 0
        sol[name] = U
    for name in ("learned", "learned fgmres"):
        e = float(torch.linalg.norm(sol[name] - sol["plain"]) / torch.linalg.norm(sol["plain"]))
        print("velocity difference %s vs plain %.2e" % (name, e), flush=True)
        assert e <= 1e-5
    pc.close()

"""GPU parity of the learned preconditioner's solver interface against the
oracle (oracle/learned.py), and what the trained network gains.

Tolerances: canonical frames 1e-12 absolute; local operator 1e-9 relative
(the FP64 matvec's own bar), 1e-6 in precision 2 (its FP32 copy of the
singular operator); P(r) 1e-5 relative to max|P(r)| (the network
runs in FP32), 2e-3 with BF16 activations against the oracle's BF16 form
(oracle/mlp.py); the inner-solve preconditioner inverts the local operator to
its 1e-12 inner tolerance (1e-9 against the oracle's operator).  Gain: after
training on solver-generated pairs from one suspension, the learned solve of
another suspension takes at most 3/4 of the diagonal-preconditioned
iterations, the learned inner solves take fewer iterations than diagonal
ones, every form reaches true relative residual 1e-6, and every velocity is
within 1e-5 of a tightly converged FP64 solve -- with FP32 and with BF16
activations.
"""
import numpy as np
import pytest
import torch

from oracle import learned as L
from oracle import ves
from ves3d_amd import implicit as I
from ves3d_amd import workloads as W

pytestmark = pytest.mark.gpu


def cu(a):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).cuda()


def rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def _ellipsoids(nv, p, vf, seed):
    return W.implicit_step_case(nv, p, vf, seed=seed)


def test_frames_and_local_matvec(hip):
    nv, p = 4, 8
    X = _ellipsoids(nv, p, 0.1, seed=31)
    st = I.Stepper(nv, p, precision=0, precond=1)
    st.set_state(cu(X))
    st.prepare()
    print("frames / local matvec nv=%d p=%d" % (nv, p), flush=True)
    F = st.frames().cpu().numpy()
    assert np.max(np.abs(F - L.frames(X, p))) < 1e-12
    orc = ves.Implicit(X, p, dt=1e-3, shear=1.0, near=False)
    x = np.random.default_rng(2).standard_normal(st.n)
    xg = cu(x)
    y = st.matvec_local(xg).cpu().numpy()
    y0 = orc.matvec(x, local=True)
    assert rel(y, y0) < 1e-9
    st.close()
    # precision 2: the local operator applies an FP32 copy of the FP64
    # singular operator (rounding 6e-8 per entry)
    st = I.Stepper(nv, p, precision=2, precond=1)
    st.set_state(cu(X))
    st.prepare()
    y2 = st.matvec_local(xg).cpu().numpy()
    print("local matvec, precision 2: %.2e" % rel(y2, y0), flush=True)
    assert 1e-12 < rel(y2, y0) < 1e-6
    st.close()


def test_inner_precond_inverts_local_operator(hip):
    """y = P_inner(x) solves A_loc y = x (every vesicle's own operator)."""
    nv, p = 3, 8
    X = _ellipsoids(nv, p, 0.1, seed=35)
    st = I.Stepper(nv, p, precision=0, precond=1, inner_max_iter=300, inner_tol=1e-12)
    st.set_state(cu(X))
    st.prepare()
    print("inner precond nv=%d p=%d" % (nv, p), flush=True)
    # a right-hand side in the local operator's range (its tension rows of
    # degree > 2p/3 vanish)
    z = cu(np.random.default_rng(3).standard_normal(st.n))
    x = st.matvec_local(z)
    y, its = st.precond_inner(x)
    assert 0 < its <= 300
    r = st.matvec_local(y) - x
    assert float(torch.linalg.norm(r) / torch.linalg.norm(x)) < 1e-10
    orc = ves.Implicit(X, p, dt=1e-3, shear=1.0, near=False)
    xo = x.cpu().numpy()
    ro = orc.matvec(y.cpu().numpy(), local=True) - xo
    assert np.linalg.norm(ro) / np.linalg.norm(xo) < 1e-9
    st.close()


def test_precond_learned_vs_oracle(hip):
    nv, p = 3, 6
    X = _ellipsoids(nv, p, 0.1, seed=32)
    nin = 4 * p * (p + 2)
    pc = I.LearnedPC(nin, 64, capacity=0, seed=5)
    rng = np.random.default_rng(6)
    params = [rng.standard_normal((64, nin)) * 0.1, rng.standard_normal(64) * 0.1,
              rng.standard_normal((64, 64)) * 0.1, rng.standard_normal(64) * 0.1,
              rng.standard_normal((nin, 64)) * 0.1, rng.standard_normal(nin) * 0.1]
    flat = np.concatenate([q.reshape(-1) for q in params]).astype(np.float32)
    pc.set_params(torch.from_numpy(flat).cuda())
    p32 = [q.astype(np.float32).astype(np.float64) for q in params]
    st = I.Stepper(nv, p, dt=1e-3, shear_rate=1.0, precision=0, precond=2)
    I.attach(st, pc, collect=False)
    st.set_state(cu(X))
    st.prepare()
    print("learned precond nv=%d p=%d" % (nv, p), flush=True)
    orc = ves.Implicit(X, p, dt=1e-3, shear=1.0, near=False)
    x = rng.standard_normal(st.n)
    xg = cu(x)
    y = st.precond_learned(xg).cpu().numpy()
    want = L.precond(orc, p32, x, L.frames(X, p))
    assert rel(y, want) < 1e-5
    # BF16 activations against the oracle's BF16 form
    pc.set_activations("bf16")
    yb = st.precond_learned(xg).cpu().numpy()
    wantb = L.precond(orc, p32, x, L.frames(X, p), bf16_ops=True)
    print("bf16 activations: vs BF16 oracle %.2e, vs FP64 %.2e" % (rel(yb, wantb), rel(yb, want)),
          flush=True)
    assert rel(yb, wantb) < 2e-3
    # the gauge: tension coefficients above the divergence filter keep D r
    keep = np.tile(L.tension_gauge(p, 2 * p // 3), (nv, 1))
    d = L.rows(orc, orc.precond(x))
    assert np.max(np.abs(L.rows(orc, y)[~keep] - d[~keep])) < 1e-13 * np.abs(d).max()
    st.close()
    pc.close()


@pytest.mark.parametrize("activations", ["fp32", "bf16"])
def test_trained_network_gain(hip, activations):
    """Trained on pairs from one suspension, evaluated on another, with FP32
    or BF16 activations."""
    nv, p = 8, 8
    common = dict(dt=1e-3, shear_rate=1.0, tol=1e-6, max_iter=300, post=False)
    nin = 4 * p * (p + 2)
    pc = I.LearnedPC(nin, 256, capacity=8192, seed=7)
    pc.set_activations(activations)
    train = I.Stepper(nv, p, precision=0, precond=1, **common)
    I.attach(train, pc, collect=True)
    for seed in (41, 42, 43):
        train.set_state(cu(_ellipsoids(nv, p, 0.2, seed)))
        train.step()
    print("training on %d samples" % pc.samples(), flush=True)
    loss = pc.train(300, 256, 1e-3)
    print("trained: loss %.3e" % loss, flush=True)
    train.close()
    X = _ellipsoids(nv, p, 0.2, seed=44)
    its = {}
    inner = {}
    U = {}
    for name, kind, tol, nin_it in (("ref", 1, 1e-11, 0), ("diag", 1, 1e-6, 0),
                                    ("learned", 2, 1e-6, 0), ("diag_inner", 1, 1e-6, 20),
                                    ("learned_inner", 2, 1e-6, 20)):
        st = I.Stepper(nv, p, precision=0, precond=kind,
                       **dict(common, tol=tol, inner_max_iter=nin_it, inner_tol=1e-3))
        if kind == 2:
            I.attach(st, pc, collect=False)
        st.set_state(cu(X))
        s = st.step()
        print(name, s["iterations"], s["inner_iters"], s["true_relres"], s["fell_back"],
              flush=True)
        assert s["status"] == 100 and s["true_relres"] < tol * 1.05
        its[name] = s["iterations"]
        inner[name] = s["inner_iters"]
        U[name] = st.get_state()[1].cpu().numpy()
        if kind == 2:
            assert s["fell_back"] == 0
        st.close()
    assert its["learned"] <= 0.75 * its["diag"], its
    # the inner solves invert each vesicle's own operator: fewer outer
    # iterations than one application of either preconditioner, and the
    # network makes the inner solves cheaper than the diagonal does
    assert its["learned_inner"] <= its["learned"] and its["diag_inner"] <= its["learned"], its
    assert inner["learned_inner"] < inner["diag_inner"], inner
    for name in ("diag", "learned", "diag_inner", "learned_inner"):
        e = np.linalg.norm(U[name] - U["ref"]) / np.linalg.norm(U["ref"])
        assert e < 1e-5, (name, e)
    pc.close()

"""The learned preconditioner's network: the numpy oracle (oracle/mlp.py) is
pinned by central finite differences in FP64 (CPU); the GPU forward, loss,
gradients and Adam step are checked against the oracle (GPU).

GPU tolerances (FP32 kernels vs the FP64 oracle, relative to the max):
  forward 1e-5, loss 1e-5, gradients 1e-4, parameters after Adam 1e-5.
"""
import numpy as np
import pytest

from oracle import mlp


def _params(nin, w, rng, scale=0.3):
    return [rng.standard_normal((w, nin)) * scale, rng.standard_normal(w) * 0.1,
            rng.standard_normal((w, w)) * scale, rng.standard_normal(w) * 0.1,
            rng.standard_normal((nin, w)) * scale, rng.standard_normal(nin) * 0.1]


def test_oracle_gradients_finite_difference():
    rng = np.random.default_rng(0)
    nin, w, B = 7, 5, 4
    P = _params(nin, w, rng)
    Z = rng.standard_normal((B, nin))
    T = rng.standard_normal((B, nin))
    loss, G = mlp.loss_and_grads(P, Z, T)
    h = 1e-6
    for k, (p, g) in enumerate(zip(P, G)):
        flat = p.reshape(-1)
        for idx in range(0, flat.size, max(1, flat.size // 6)):
            old = flat[idx]
            flat[idx] = old + h
            lp, _ = mlp.loss_and_grads(P, Z, T)
            flat[idx] = old - h
            lm, _ = mlp.loss_and_grads(P, Z, T)
            flat[idx] = old
            fd = (lp - lm) / (2 * h)
            assert abs(fd - g.reshape(-1)[idx]) < 1e-7 * max(1.0, abs(fd)), (k, idx)


def test_oracle_gelu_matches_definition():
    x = np.linspace(-4, 4, 33)
    from math import erf, sqrt
    exact = np.array([0.5 * v * (1 + erf(v / sqrt(2))) for v in x])
    assert np.max(np.abs(mlp.gelu(x) - exact)) < 2e-3   # tanh form vs erf form
    h = 1e-6
    assert np.allclose(mlp.gelu_grad(x), (mlp.gelu(x + h) - mlp.gelu(x - h)) / (2 * h), atol=1e-8)


@pytest.mark.gpu
def test_gpu_network_vs_oracle(hip):
    import torch
    from ves3d_amd.implicit import LearnedPC
    rng = np.random.default_rng(1)
    nin, w, B = 96, 64, 37
    pc = LearnedPC(nin, w, capacity=0, seed=3)
    P = _params(nin, w, rng, scale=0.2)
    flat = np.concatenate([p.reshape(-1) for p in P]).astype(np.float32)
    assert flat.size == pc.num_params
    pc.set_params(torch.from_numpy(flat).cuda())
    P32 = [p.astype(np.float32).astype(np.float64) for p in P]
    Z = rng.standard_normal((B, nin)).astype(np.float32)
    T = rng.standard_normal((B, nin)).astype(np.float32)
    Y = pc.forward(torch.from_numpy(Z).cuda()).cpu().numpy()
    Y0, _ = mlp.forward(P32, Z.astype(np.float64))
    assert np.max(np.abs(Y - Y0)) / np.max(np.abs(Y0)) < 1e-5
    loss, g = pc.loss_grad(torch.from_numpy(Z).cuda(), torch.from_numpy(T).cuda())
    loss0, G0 = mlp.loss_and_grads(P32, Z.astype(np.float64), T.astype(np.float64))
    assert abs(loss - loss0) / loss0 < 1e-5
    g = g.cpu().numpy()
    for (off, size, shape), g0 in zip(pc.layout(), G0):
        gg = g[off:off + size].reshape(shape)
        assert np.max(np.abs(gg - g0)) / np.max(np.abs(g0)) < 1e-4
    pc.adam(1e-3)
    newp = pc.get_params().cpu().numpy()
    zero = [np.zeros_like(p) for p in P32]
    P1, _, _ = mlp.adam_step(P32, G0, zero, zero, 1, 1e-3)
    exp = np.concatenate([p.reshape(-1) for p in P1])
    assert np.max(np.abs(newp - exp)) / np.max(np.abs(exp)) < 1e-5
    # with decoupled weight decay, from the same parameters (set_params also
    # resets the optimizer state)
    pc.set_params(torch.from_numpy(flat).cuda())
    pc.set_weight_decay(0.1)
    pc.loss_grad(torch.from_numpy(Z).cuda(), torch.from_numpy(T).cuda())
    pc.adam(1e-3)
    newp = pc.get_params().cpu().numpy()
    P1, _, _ = mlp.adam_step(P32, G0, zero, zero, 1, 1e-3, wd=0.1)
    exp = np.concatenate([p.reshape(-1) for p in P1])
    assert np.max(np.abs(newp - exp)) / np.max(np.abs(exp)) < 1e-5
    pc.close()

"""The learned preconditioner's network: the numpy oracle (oracle/mlp.py) is
pinned by central finite differences in FP64 (CPU); the GPU forward, loss,
gradients and Adam step are checked against the oracle (GPU).

GPU tolerances (FP32 kernels vs the FP64 oracle, relative to the max):
  forward 1e-5, loss 1e-5, gradients 1e-4, parameters after Adam 1e-5.
BF16 activations (BF16 MFMA operands, FP32 accumulation) against the
oracle's BF16 restatement (operands rounded, products exact): forward 2e-3,
gradients 5e-3 (a pre-activation that lands within FP32 rounding of a
bfloat16 tie can round the other way: one ulp, 2^-8, of one activation);
against the unrounded FP64 network: forward and gradients 5e-2, and the
forward differs by more than 1e-4 (the BF16 path is the one that ran).
The oracle's rounding is pinned to torch's float32 -> bfloat16 conversion.
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


def test_oracle_bf16_rounding_matches_torch():
    import torch
    rng = np.random.default_rng(2)
    x = np.concatenate([rng.standard_normal(20000) * 10.0 ** rng.integers(-30, 30, 20000),
                        [0.0, -0.0, 1.0, 1.00390625, 1.005859375, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8,
                         65504.0, 3.0e38]])
    want = torch.from_numpy(x.astype(np.float32)).to(torch.bfloat16).to(torch.float64).numpy()
    got = mlp.bf16(x)
    assert np.array_equal(got, want)
    # ties go to even: 1 + 2^-8 lies halfway between 1 and 1 + 2^-7
    assert mlp.bf16(1.0 + 2 ** -8) == 1.0 and mlp.bf16(1.0 + 3 * 2 ** -8) == 1.0 + 2 ** -6


def test_oracle_bf16_form_reduces_to_fp64_when_exact():
    """Operands that are bfloat16 already: the BF16 form equals the FP64
    network (the rounding is the only difference)."""
    rng = np.random.default_rng(3)
    nin, w, B = 16, 8, 5
    P = [mlp.bf16(p) for p in _params(nin, w, rng)]
    Z, T = mlp.bf16(rng.standard_normal((B, nin))), rng.standard_normal((B, nin))
    Y0, c0 = mlp.forward(P, Z)
    Y1, c1 = mlp.forward(P, Z, bf16_ops=True)
    # the activations differ by their rounding, and only by it
    assert np.array_equal(c1[2], mlp.bf16(c0[2]))
    assert np.max(np.abs(Y1 - Y0)) < 1e-2 * np.max(np.abs(Y0))
    l0, G0 = mlp.loss_and_grads(P, Z, T)
    l1, G1 = mlp.loss_and_grads(P, Z, T, bf16_ops=True)
    assert abs(l1 - l0) < 1e-2 * l0
    for g0, g1 in zip(G0, G1):
        assert np.max(np.abs(g1 - g0)) < 5e-2 * np.max(np.abs(g0))


def _rel(a, b):
    return np.max(np.abs(a - b)) / np.max(np.abs(b))


@pytest.mark.gpu
@pytest.mark.parametrize("nin,w,B", [(96, 64, 37), (1152, 256, 300)])
def test_gpu_network_bf16_vs_oracle(hip, nin, w, B):
    import torch
    from ves3d_amd.implicit import LearnedPC
    rng = np.random.default_rng(4)
    pc = LearnedPC(nin, w, capacity=0, seed=3)
    P = _params(nin, w, rng, scale=1.0 / np.sqrt(nin))
    P[2] *= np.sqrt(nin / w)
    P[4] *= np.sqrt(nin / w)
    flat = np.concatenate([p.reshape(-1) for p in P]).astype(np.float32)
    pc.set_params(torch.from_numpy(flat).cuda())
    pc.set_activations("bf16")
    P32 = [p.astype(np.float32).astype(np.float64) for p in P]
    Z = rng.standard_normal((B, nin)).astype(np.float32)
    T = rng.standard_normal((B, nin)).astype(np.float32)
    Zg, Tg = torch.from_numpy(Z).cuda(), torch.from_numpy(T).cuda()
    print("bf16 network nin=%d width=%d batch=%d" % (nin, w, B), flush=True)
    Y = pc.forward(Zg).cpu().numpy()
    Yb, _ = mlp.forward(P32, Z.astype(np.float64), bf16_ops=True)
    Y0, _ = mlp.forward(P32, Z.astype(np.float64))
    print("forward vs BF16 oracle %.2e, vs FP64 %.2e" % (_rel(Y, Yb), _rel(Y, Y0)), flush=True)
    assert _rel(Y, Yb) < 2e-3
    assert 1e-4 < _rel(Y, Y0) < 5e-2
    loss, g = pc.loss_grad(Zg, Tg)
    lb, Gb = mlp.loss_and_grads(P32, Z.astype(np.float64), T.astype(np.float64), bf16_ops=True)
    l0, G0 = mlp.loss_and_grads(P32, Z.astype(np.float64), T.astype(np.float64))
    assert abs(loss - lb) / lb < 1e-4
    g = g.cpu().numpy()
    for (off, size, shape), gb, g0 in zip(pc.layout(), Gb, G0):
        gg = g[off:off + size].reshape(shape)
        print("  grad %s: vs BF16 oracle %.2e, vs FP64 %.2e" % (shape, _rel(gg, gb), _rel(gg, g0)),
              flush=True)
        assert _rel(gg, gb) < 5e-3 and _rel(gg, g0) < 5e-2
    # an Adam step refreshes the BF16 weights from the FP32 master copy
    pc.adam(1e-3)
    newp = pc.get_params().cpu().numpy()
    zero = [np.zeros_like(p) for p in P32]
    P1, _, _ = mlp.adam_step(P32, Gb, zero, zero, 1, 1e-3)
    P1 = [np.asarray(a) for a in P1]
    # Adam's first step moves each parameter by about lr sign(g): exact up
    # to the gradients whose sign the rounding can flip (|g| near 0)
    exp = np.concatenate([p.reshape(-1) for p in P1])
    assert np.mean(np.abs(newp - exp) < 1e-6) > 0.99
    Y1 = pc.forward(Zg).cpu().numpy()
    Y1b, _ = mlp.forward([p.astype(np.float32).astype(np.float64) for p in
                          [newp[o:o + n].reshape(sh) for o, n, sh in pc.layout()]],
                         Z.astype(np.float64), bf16_ops=True)
    assert _rel(Y1, Y1b) < 2e-3
    # back to FP32 activations: the FP32 result
    pc.set_activations("fp32")
    Y2 = pc.forward(Zg).cpu().numpy()
    Y2o, _ = mlp.forward([newp[o:o + n].reshape(sh).astype(np.float64) for o, n, sh in pc.layout()],
                         Z.astype(np.float64))
    assert _rel(Y2, Y2o) < 1e-5
    pc.close()


def test_stored_network_files():
    """The committed network: FP32 parameters of the p = 16, width 256
    layout, finite, and a record of the 200 plain-solve steps it was trained
    on (read with loaders that execute nothing)."""
    import json
    import os
    from ves3d_amd import training as T
    base = T.weights_path(16, 256)
    flat = np.load(base + ".npy", allow_pickle=False)
    nin, w = 4 * 16 * 18, 256
    assert flat.dtype == np.float32 and flat.shape == (w * nin * 2 + w * w + 2 * w + nin,)
    assert np.all(np.isfinite(flat))
    meta = json.load(open(base + ".json"))
    assert meta["p"] == 16 and meta["width"] == w and meta["nin"] == nin
    assert meta["plain_solve_steps"] == 200 and len(meta["plain_iterations"]) == 200
    assert meta["activations"] == "bf16" and meta["samples"] > 500000
    assert os.path.basename(base) == "lpc_p16_w256"

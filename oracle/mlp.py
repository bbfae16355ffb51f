"""FP64 numpy oracle of the learned SH-space preconditioner network.

TEST INFRASTRUCTURE ONLY: the checker for the GPU forward / backward / Adam
kernels (ves3d_amd/csrc/learned_pc.hip).  The network has no counterpart in
the reference; its role is the preconditioner slot of
ParallelLinSolver::UpdatePrecond / ImplicitPrecond
(src/InterfacialVelocity.cc:1054-1109).

Network (per vesicle, rows of Z):  H1 = gelu(Z W1^T + b1),
H2 = gelu(H1 W2^T + b2),  Y = H2 W3^T + b3;  loss = mean((Y - T)^2) / 2
over all entries.  gelu is the tanh form
  gelu(x) = x/2 (1 + tanh(sqrt(2/pi)(x + 0.044715 x^3))).

bf16_ops=True restates the BF16-activation form: both operands of every
product (Z, H1, H2, the weights; dY, dA2, dA1 in the backward pass) rounded
to the nearest bfloat16, the products and everything else exact (FP64).
"""
import numpy as np

C0 = np.sqrt(2.0 / np.pi)
C1 = 0.044715


def gelu(x):
    return 0.5 * x * (1.0 + np.tanh(C0 * (x + C1 * x ** 3)))


def gelu_grad(x):
    u = C0 * (x + C1 * x ** 3)
    t = np.tanh(u)
    return 0.5 * (1.0 + t) + 0.5 * x * (1.0 - t * t) * C0 * (1.0 + 3.0 * C1 * x * x)


def bf16(x):
    """The nearest bfloat16 (ties to even) of float32(x), as float64: the
    float32 bit pattern rounded at bit 16 (bfloat16 is float32's sign,
    exponent and top 7 fraction bits)."""
    f = np.ascontiguousarray(x, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def _rounding(bf16_ops):
    return bf16 if bf16_ops else (lambda a: a)


def forward(params, Z, bf16_ops=False):
    r = _rounding(bf16_ops)
    W1, b1, W2, b2, W3, b3 = params
    Z = r(Z)
    A1 = Z @ r(W1).T + b1
    H1 = r(gelu(A1))
    A2 = H1 @ r(W2).T + b2
    H2 = r(gelu(A2))
    Y = H2 @ r(W3).T + b3
    return Y, (Z, A1, H1, A2, H2)


def loss_and_grads(params, Z, T, bf16_ops=False):
    r = _rounding(bf16_ops)
    W1, b1, W2, b2, W3, b3 = params
    Y, (Z, A1, H1, A2, H2) = forward(params, Z, bf16_ops)
    n = Y.size
    D = Y - T
    loss = 0.5 * np.sum(D * D) / n
    dY = D / n
    gW3 = r(dY).T @ H2
    gb3 = dY.sum(axis=0)
    dH2 = r(dY) @ r(W3)
    dA2 = dH2 * gelu_grad(A2)
    gW2 = r(dA2).T @ H1
    gb2 = dA2.sum(axis=0)
    dH1 = r(dA2) @ r(W2)
    dA1 = dH1 * gelu_grad(A1)
    gW1 = r(dA1).T @ Z
    gb1 = dA1.sum(axis=0)
    return loss, (gW1, gb1, gW2, gb2, gW3, gb3)


def adam_step(params, grads, m, v, t, lr, b1=0.9, b2=0.999, eps=1e-8, wd=0.0):
    """One Adam update (Kingma & Ba 2015, Algorithm 1) on every tensor, with
    decoupled weight decay wd (Loshchilov & Hutter 2019)."""
    out_p, out_m, out_v = [], [], []
    for p, g, mm, vv in zip(params, grads, m, v):
        mm = b1 * mm + (1 - b1) * g
        vv = b2 * vv + (1 - b2) * g * g
        mh = mm / (1 - b1 ** t)
        vh = vv / (1 - b2 ** t)
        out_p.append(p - lr * (mh / (np.sqrt(vh) + eps) + wd * p))
        out_m.append(mm)
        out_v.append(vv)
    return out_p, out_m, out_v

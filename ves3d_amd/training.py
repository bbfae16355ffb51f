"""Training of the learned preconditioner on pairs that plain solves
generate (host orchestration; the recording, the network's forward and
backward passes and Adam run in the library's own kernels).

The pairs are (D A_loc e, e) for every Krylov vector e of the plain
(FP64, diagonal-preconditioned) solves of training suspensions of the same
shape family, recorded by a stepper with the network attached in collect
mode.
"""
import time

import numpy as np

from . import implicit as I
from . import workloads as W


def train_learned_pc(nves, p, volfracs, common, width=256, capacity=200000, epochs=20,
                     batch=512, lr=1e-3, steps_per_suspension=1, seed=7, log=None):
    """Records pairs from plain solves of one suspension per volume fraction
    (perturbed ellipsoids, `nves` vesicles, seeds 1000 + k), each run for
    `steps_per_suspension` steps, then trains the network.

    common: the Stepper keywords of the solves (dt, shear_rate, tol, ...).
    Returns (pc, info) with info = dict(collect_s, train_s, samples, loss,
    steps)."""
    import torch
    nin = 4 * p * (p + 2)
    pc = I.LearnedPC(nin, width, capacity=capacity, seed=seed)
    t0 = time.perf_counter()
    trainer = I.Stepper(nves, p, precision=0, precond=1, **common)
    I.attach(trainer, pc, collect=True)
    steps = 0
    for k, vf in enumerate(volfracs):
        X = W.implicit_step_case(nv=nves, p=p, volfrac=float(vf), seed=1000 + k)
        trainer.set_state(torch.from_numpy(X).cuda())
        for j in range(steps_per_suspension):
            s = trainer.step()
            steps += 1
            if log:
                log("training suspension %d (%d ves, %.0f%%) step %d: %d iters, %d samples" %
                    (k, nves, 100 * vf, j, s["iterations"], pc.samples()))
    trainer.close()
    collect_s = time.perf_counter() - t0
    t0 = time.perf_counter()
    loss = pc.train(epochs, batch, lr)
    train_s = time.perf_counter() - t0
    return pc, dict(collect_s=collect_s, train_s=train_s, samples=pc.samples(), loss=loss,
                    steps=steps)

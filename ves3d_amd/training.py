"""Training of the learned preconditioner on pairs that plain solves
generate (host orchestration; the recording, the network's forward and
backward passes and Adam run in the library's own kernels).

The pairs are (D A_loc e, e) for every Krylov vector e of the plain
(FP64, diagonal-preconditioned) solves of training suspensions of the same
shape family, recorded by a stepper with the network attached in collect
mode.

Run as a module, it trains the stored network the bench loads: the first
200 plain-solve steps of 256-vesicle suspensions at 10, 15, 20 and 25%
volume fraction (50 steps each), p = 16, then writes the FP32 parameters
(`<out>.npy`) and how they were made (`<out>.json`):

    python -m ves3d_amd.training --out gpurun_out/lpc_p16_w256
"""
import argparse
import json
import os
import time

import numpy as np

from . import implicit as I
from . import workloads as W

WEIGHTS_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "weights")


def train_learned_pc(nves, p, volfracs, common, width=256, capacity=200000, epochs=20,
                     batch=512, lr=1e-3, steps_per_suspension=1, seed=7, log=None,
                     activations="fp32"):
    """Records pairs from plain solves of one suspension per volume fraction
    (perturbed ellipsoids, `nves` vesicles, seeds 1000 + k), each run for
    `steps_per_suspension` steps, then trains the network.

    common: the Stepper keywords of the solves (dt, shear_rate, tol, ...).
    Returns (pc, info) with info = dict(collect_s, train_s, samples, loss,
    steps, iterations)."""
    import torch
    nin = 4 * p * (p + 2)
    pc = I.LearnedPC(nin, width, capacity=capacity, seed=seed)
    pc.set_activations(activations)
    t0 = time.perf_counter()
    trainer = I.Stepper(nves, p, precision=0, precond=1, **common)
    I.attach(trainer, pc, collect=True)
    steps, iters = 0, []
    for k, vf in enumerate(volfracs):
        X = W.implicit_step_case(nv=nves, p=p, volfrac=float(vf), seed=1000 + k)
        trainer.set_state(torch.from_numpy(X).cuda())
        for j in range(steps_per_suspension):
            s = trainer.step()
            steps += 1
            iters.append(int(s["iterations"]))
            if log and (j < 2 or j % 10 == 9 or j == steps_per_suspension - 1):
                log("training suspension %d (%d ves, %.0f%%) step %d: %d iters, %d samples" %
                    (k, nves, 100 * vf, j, s["iterations"], pc.samples()))
    trainer.close()
    collect_s = time.perf_counter() - t0
    t0 = time.perf_counter()
    loss = pc.train(epochs, batch, lr)
    train_s = time.perf_counter() - t0
    return pc, dict(collect_s=collect_s, train_s=train_s, samples=pc.samples(), loss=loss,
                    steps=steps, iterations=iters)


def weights_path(p, width):
    """The stored network for SH order p and hidden width (without suffix)."""
    return os.path.join(WEIGHTS_DIR, "lpc_p%d_w%d" % (p, width))


def load_stored(p, width, activations="bf16", path=None):
    """A LearnedPC with the stored parameters and their provenance, or
    (None, None) when none are stored for (p, width)."""
    import torch
    base = path or weights_path(p, width)
    if not os.path.exists(base + ".npy"):
        return None, None
    flat = np.load(base + ".npy", allow_pickle=False)
    with open(base + ".json") as f:
        info = json.load(f)
    nin = 4 * p * (p + 2)
    pc = I.LearnedPC(nin, width, capacity=0, seed=info.get("seed", 7))
    if flat.shape != (pc.num_params,) or flat.dtype != np.float32:
        raise ValueError("%s.npy: %s %s, expected (%d,) float32" %
                         (base, flat.shape, flat.dtype, pc.num_params))
    pc.set_params(torch.from_numpy(flat).cuda())
    pc.set_activations(activations)
    return pc, info


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--out", required=True, help="output path without suffix")
    ap.add_argument("--nves", type=int, default=256)
    ap.add_argument("--p", type=int, default=16)
    ap.add_argument("--steps", type=int, default=200, help="plain-solve steps in all")
    ap.add_argument("--fracs", default="0.10,0.15,0.20,0.25")
    ap.add_argument("--width", type=int, default=256)
    ap.add_argument("--capacity", type=int, default=1400000)
    ap.add_argument("--epochs", default="20", help="comma-separated: one network per count")
    ap.add_argument("--batch", type=int, default=1024)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--activations", default="bf16", choices=["fp32", "bf16"])
    ap.add_argument("--seed", type=int, default=7)
    args = ap.parse_args()
    import torch

    def log(msg):
        print(msg, flush=True)

    fracs = [float(v) for v in args.fracs.split(",")]
    per = max(1, args.steps // len(fracs))
    common = dict(dt=1e-3, shear_rate=1.0, tol=1e-6, max_iter=300, post=False, r_near=0.1)
    epochs = [int(e) for e in args.epochs.split(",")]
    pc, info = train_learned_pc(args.nves, args.p, fracs, common, width=args.width,
                                capacity=args.capacity, epochs=epochs[0], batch=args.batch,
                                lr=args.lr, steps_per_suspension=per, seed=args.seed, log=log,
                                activations=args.activations)
    log("collected %d samples from %d steps in %.1f s" %
        (info["samples"], info["steps"], info["collect_s"]))
    os.makedirs(os.path.dirname(os.path.abspath(args.out)) or ".", exist_ok=True)
    for k, ep in enumerate(epochs):
        if k > 0:
            # a fresh network of the same seed on the same samples
            t0 = time.perf_counter()
            flat0 = I.LearnedPC(pc.nin, args.width, capacity=0, seed=args.seed).get_params()
            pc.set_params(flat0)
            info["loss"] = pc.train(ep, args.batch, args.lr)
            info["train_s"] = time.perf_counter() - t0
        loss, tms = pc.dataset_loss()
        out = args.out if len(epochs) == 1 else "%s_e%d" % (args.out, ep)
        flat = pc.get_params().cpu().numpy().astype(np.float32)
        np.save(out + ".npy", flat, allow_pickle=False)
        meta = dict(
            p=args.p, width=args.width, nin=pc.nin, seed=args.seed,
            activations=args.activations, epochs=ep, batch=args.batch, lr=args.lr,
            training_suspensions="%d x %d perturbed ellipsoids (1, 1, 0.6 +- 0.05) at %s "
                                 "volume fraction, seeds 1000..%d" % (
                                     len(fracs), args.nves,
                                     "/".join("%.0f%%" % (100 * f) for f in fracs),
                                     999 + len(fracs)),
            plain_solve_steps=info["steps"], steps_per_suspension=per,
            step="FP64 plain solve, diagonal preconditioner, BiCGStab 1e-6, shear 1.0, "
                 "dt 1e-3, r_near 0.1",
            plain_iterations=info["iterations"], samples=int(info["samples"]),
            samples_kept=int(min(info["samples"], args.capacity)),
            heldout_loss=info["loss"], dataset_rel_fit=float(np.sqrt(loss / (0.5 * tms))),
            collect_seconds=info["collect_s"], train_seconds=info["train_s"],
            command="python -m ves3d_amd.training " + " ".join(
                "--%s %s" % (k2.replace("_", "-"), v) for k2, v in sorted(vars(args).items())))
        with open(out + ".json", "w") as f:
            json.dump(meta, f, indent=1)
        log("epochs %d: held-out loss %.3e, relative fit %.3f, %.1f s -> %s" %
            (ep, info["loss"], meta["dataset_rel_fit"], info["train_s"], out))
    pc.close()
    torch.cuda.synchronize()


if __name__ == "__main__":
    main()

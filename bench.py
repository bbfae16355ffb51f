"""Benchmark driver.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME]

Prints ONE JSON line (rank 0) with the measured metric, a roofline record for
the dominant kernel (timed with HIP events on the launch stream inside the
timed region) and a CPU baseline measured on a bounded sample on the host.

Workloads
  far : the far-field direct Stokes SL+DL sum of 512 unit-sphere vesicles at
        p=16 (278,528 points; own-vesicle pairs excluded), FP64 and
        FP32-compensated.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP64_TFLOPS = 78.6    # MI355X FP64 vector / matrix
PEAK_FP32_TFLOPS = 157.3   # MI355X FP32 vector (published; counts packed FP32)
SL_FLOP_PER_PAIR = 28      # src/CPUKernels.cc:1008
DL_FLOP_PER_PAIR = 31      # src/PVFMMInterface.cc:365 (no Newton steps)


def _cpu_threads():
    return int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))


def bench_far(args):
    import torch
    import ves3d_amd
    from ves3d_amd import stokes as K
    from ves3d_amd import workloads as W

    ves3d_amd.init(0)
    pos, sld, dld, npv = W.stokes_direct_sum_case(nv=512, p=16)
    n = pos.shape[0]
    g = lambda a: torch.from_numpy(a).cuda()
    gp, gs, gd = g(pos), g(sld), g(dld)
    out = torch.zeros((n, 3), dtype=torch.float64, device="cuda")
    mixed = args.dtype == "f32c"
    for _ in range(args.warmup):
        K.direct_eval(gp, gs, gd, gp, out, npv, npv, mixed=mixed)
    torch.cuda.synchronize()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(args.steps)]
    t0 = time.perf_counter()
    for e0, e1 in ev:
        e0.record()
        K.direct_eval(gp, gs, gd, gp, out, npv, npv, mixed=mixed)
        e1.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    kms = float(np.mean([a.elapsed_time(b) for a, b in ev]))
    pairs = n * (n - npv)
    flop = pairs * (SL_FLOP_PER_PAIR + DL_FLOP_PER_PAIR)
    ms = wall / args.steps * 1e3
    achieved = flop / (kms * 1e-3) / 1e12
    peak = PEAK_FP32_TFLOPS if mixed else PEAK_FP64_TFLOPS

    cpu = None
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import oracle_lib as OL
        orc = OL.load()
        m = 256
        idx = np.arange(0, n, n // m)[:m]
        tp = np.ascontiguousarray(pos[idx])
        r = np.zeros((m, 3))
        t0 = time.perf_counter()
        orc.orc_far_field(OL.p(pos), OL.p(sld), OL.p(dld), OL.cs(n), OL.p(tp), OL.p(r), OL.cs(m),
                          OL.cs(0), OL.cs(0))
        dt = time.perf_counter() - t0
        cpu = {"value": dt / (m * n) * pairs * 1e3, "unit": "ms per far-field sum (extrapolated)",
               "cores": _cpu_threads(), "kind": "port",
               "sample": "oracle FP64 SL+DL sum for %d of %d targets over all %d sources, scaled by pair count"
                         % (m, n, n)}
    return {
        "metric": "ms per far-field Stokes SL+DL direct sum @512 ves p=16",
        "value": ms, "unit": "ms", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms, "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32" if mixed else "f64", "data": "synthetic",
        "config": {"workload": "far-field direct sum, 512 unit spheres p=16, 15% volume fraction",
                   "points": n, "pairs": pairs, "precision": args.dtype},
        "roofline": {"bound": "valu", "achieved": achieved, "peak": peak, "unit": "TFLOP/s",
                     "frac": achieved / peak, "traffic": None,
                     "kernel": "k_far_%s<SL,DL>" % ("f32c" if mixed else "f64"),
                     "kernel_ms": kms, "flop_per_launch": flop},
        "cpu_baseline": cpu,
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="far", choices=["far"])
    ap.add_argument("--dtype", default="f64", choices=["f64", "f32c"])
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU; --gpus %d is not supported "
              "(the path shards by vesicle, see DESIGN.md)" % args.gpus, file=sys.stderr)
        sys.exit(2)
    res = bench_far(args)
    print(json.dumps(res))


if __name__ == "__main__":
    main()

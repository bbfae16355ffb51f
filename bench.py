"""Benchmark driver.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME]

Prints ONE JSON line with the measured metric, a roofline record for the
dominant kernel (timed with HIP events on its launch stream inside the timed
region) and a CPU baseline measured on a bounded sample on the host.

Workloads
  step (default): the full implicit time step of 512 perturbed ellipsoids
        (semi-axes 1, 1, 0.6 +- 0.05, random orientations) at p=16, 20% volume
        fraction, shear rate 1.0, dt = 1e-3, the outer solve to relative
        residual 1e-6 (flexible GMRES, the method of the reference's implicit
        solve; --solver bicgstab for BiCGStab with a true-residual check).
        Two forms run back to back from the same state:
          plain  : FP64 everywhere, diagonal spectral preconditioner;
          learned: FP32-compensated far field (and, with --mixed-precision 1,
                   the FP32 singular operator) and the learned SH-space
                   preconditioner, trained beforehand (untimed) on
                   (A_loc e, e) pairs recorded during plain solves of separate
                   256-vesicle training suspensions at 10-25% volume fraction.
                   Each preconditioner application is an inner BiCGStab on
                   the vesicles' own operators (no far field) preconditioned
                   by the network (--inner-max-iter, --inner-tol; 0 applies
                   the network once).  The same inner solves with the
                   diagonal preconditioner inside run afterwards, reported
                   under `inner_solve`.
        The velocity error of both forms' first step is measured against an
        FP64 solve of the same system converged to relative residual 1e-10.
        `value` is the learned + mixed form's ms per step; the plain form's is
        `ms_per_step_plain`.
  far : the far-field direct Stokes SL+DL sum alone, 512 unit spheres at
        p=16 (278,528 points; own-vesicle pairs excluded).
  matvec: K Galerkin matvecs of the step workload (--dtype f32c: the mixed
        form) after a 512 MiB device copy that calibrates HBM counters.
  probe: the VALU rate the GPU sustains (independent FMA chains, FP32
        non-packed and FP64), the attainable ceiling of the far-field kernels.
  stages: per-stage rooflines outside the far field: the batched singular
        build (SL + DL, 2048 vesicles at p=16 and 512 at p=24, FP64 and FP32
        transforms, with the FP32 error per output), the self-apply GEMV and
        the SH transforms of the 512-vesicle step.
  build: one singular build (SL + DL) of --nves vesicles at --p, FP64 and
        FP32 transforms (a short program for hardware-counter runs).
  trajectory: the step workload for --traj-steps steps (default 100) in the
        plain and the learned + mixed form from the same state: every step's
        time, iterations, fallbacks and true residual, and the positions'
        drift between the two trajectories.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "ms per implicit step @512 ves p=16 (learned-PC+mixed vs plain) at 1e-6"
PEAK_FP64_TFLOPS = 78.6    # MI355X FP64 vector / matrix
PEAK_FP32_TFLOPS = 157.3   # MI355X FP32 vector (published; counts packed FP32)
PEAK_FP32_NONPACKED_TFLOPS = 78.6   # the same without packed FP32 (disabled in this build)
SL_FLOP_PER_PAIR = 28      # src/CPUKernels.cc:1008
# HBM bytes per launch of k_far_groups_f32c at 512 vesicles, p = 16, from two
# rocprofv3 --pmc runs (FETCH_SIZE x 2.0 + WRITE_SIZE x 1.0, the factors a
# 512 MiB device copy gives the two counters in the same runs), keyed by the
# (far_coarse, far_coarse2) setting they were measured at
FAR32_TRAFFIC = {(3.0, 6.0): (2.0 * 389723.9 * 1024 + 52224.0 * 1024,
                              "profiles/k_far_groups_f32c_hbm_counters_512ves_p16.md")}
FAR_KINDS = {"lds": 0, "scalar": 1, "scalar2": 2}
FAR32_NAMES = {"lds": "k_far_f32c<SL>", "scalar": "k_far_groups_f32c", "scalar2": "k_far_groups_f32c2"}
DL_FLOP_PER_PAIR = 31      # src/PVFMMInterface.cc:365 (no Newton steps)


def _cpu_threads():
    return int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))


def _log(*a):
    print(*a, file=sys.stderr, flush=True)


# ---------------------------------------------------------------- step
def _train_pc(args, p, common):
    """The learned preconditioner, untimed.  By default the stored network
    (ves3d_amd/weights, trained by `python -m ves3d_amd.training` on the
    first 200 plain-solve steps of 256-vesicle suspensions at 10-25% volume
    fraction); with --pc-weights train, or when none is stored for (p,
    width), trained in this run on pairs that plain solves of
    args.train_steps separate training suspensions generate (args.train_nves
    vesicles, 10% .. 25%).  Returns (pc, record) with the record of how the
    network was made."""
    from ves3d_amd import training as T
    if args.pc_weights != "train":
        pc, meta = T.load_stored(p, args.width, args.activations,
                                 None if args.pc_weights == "stored" else args.pc_weights)
        if pc is not None:
            _log("stored network: %d plain-solve steps, %d samples, held-out loss %.3e" %
                 (meta["plain_solve_steps"], meta["samples"], meta["heldout_loss"]))
            rec = {"source": "stored (%s)" % os.path.relpath(
                       (None if args.pc_weights == "stored" else args.pc_weights)
                       or T.weights_path(p, args.width), os.path.dirname(os.path.abspath(__file__))),
                   "activations": args.activations}
            rec.update({k: meta[k] for k in (
                "training_suspensions", "plain_solve_steps", "step", "samples", "samples_kept",
                "epochs", "batch", "heldout_loss", "dataset_rel_fit", "collect_seconds",
                "train_seconds", "activations", "command") if k in meta})
            rec["trained_with_activations"] = meta.get("activations")
            rec["activations"] = args.activations
            return pc, rec
        if args.pc_weights != "stored":
            raise FileNotFoundError(args.pc_weights + ".npy")
        _log("no stored network for p=%d width=%d: training in this run" % (p, args.width))
    fracs = np.linspace(0.10, 0.25, max(1, args.train_steps))
    pc, info = T.train_learned_pc(args.train_nves, p, fracs, common, width=args.width,
                                  capacity=args.capacity, epochs=args.epochs, lr=args.lr,
                                  log=_log, activations=args.activations)
    _log("trained on %d samples: loss %.3e in %.1f s" % (info["samples"], info["loss"],
                                                         info["train_s"]))
    return pc, {"source": "trained in this run", "activations": args.activations,
                "training_suspensions": "%d x %d ves at %s%% volume fraction" % (
                    len(fracs), args.train_nves, "/".join("%.0f" % (100 * f) for f in fracs)),
                "plain_solve_steps": info["steps"], "samples": int(info["samples"]),
                "epochs": args.epochs, "heldout_loss": info["loss"],
                "collect_seconds": info["collect_s"], "train_seconds": info["train_s"]}


def _network_bf16_error(I, nv, p, Xg, kw, pc):
    """P(b) of the learned preconditioner with BF16 against FP32 activations
    for the first step's right-hand side b: relative to |P(b)| and to the
    network's own correction |P(b) - D b|."""
    st = I.Stepper(nv, p, **kw)
    I.attach(st, pc, collect=False)
    st.set_state(Xg)
    st.prepare()
    b = st.rhs()
    act = pc.activations
    out = {}
    for kind in ("bf16", "fp32"):
        pc.set_activations(kind)
        out[kind] = st.precond_learned(b).double()
    d = st.precond(b).double()
    pc.set_activations(act)
    st.close()
    diff = (out["bf16"] - out["fp32"]).norm().item()
    return {"rel_to_P": diff / out["fp32"].norm().item(),
            "rel_to_correction": diff / (out["fp32"] - d).norm().item()}


def _run_form(I, nv, p, X0g, kw, pc, nsteps, name, check=None, keep_first=False):
    """Runs one form of the step for nsteps steps from X0g.  check: a
    precision-0 stepper with the exact far field that judges every step's
    solution by the FP64 operator (untimed: the residual is computed between
    steps, outside the step's own clock).  Returns (stats list, FP64 relres
    list, first-step (U, sigma) if keep_first)."""
    import torch
    st = I.Stepper(nv, p, **kw)
    if pc is not None:
        I.attach(st, pc, collect=False)
    st.set_state(X0g)
    stats, r64, first = [], [], None
    for k in range(nsteps):
        Xb = st.get_state()[0].clone() if check is not None else None
        s = st.step()
        _, U, S = st.get_state()
        if check is not None:
            check.set_state(Xb, U, S)
            s["fp64_relres"] = check.residual()
            r64.append(s["fp64_relres"])
        if keep_first and k == 0:
            first = (U.clone(), S.clone())
        _log("%s step %d: %.1f ms, %d iters (%d inner), relres %.2e true %.2e%s, far %.1f ms "
             "(%d + %d two-density), near %.1f ms (%d pairs), fell_back %d" %
             (name, k, s["ms_total"], s["iterations"], s["inner_iters"], s["relres"],
              s["true_relres"], (" fp64 %.2e" % s["fp64_relres"]) if check is not None else "",
              s["ms_far"] + s["ms_far2"], s["far_calls"], s["far2_calls"], s["ms_near"],
              s["near_pairs"], s["fell_back"]))
        stats.append(s)
    st.close()
    return stats, r64, first


def _mean(stats, key):
    return float(np.mean([s[key] for s in stats]))


def _init(args):
    """Device 0 and the process-wide kernel choices of the run."""
    import ves3d_amd
    ves3d_amd.init(0)
    ves3d_amd.call("ves3d_hip_set_far_kernel", FAR_KINDS[args.far_kernel])
    ves3d_amd.call("ves3d_hip_set_pole_kernel", 1 if args.pole_kernel == "fused" else 0)


def bench_step(args):
    import torch
    import ves3d_amd
    from ves3d_amd import implicit as I
    from ves3d_amd import workloads as W

    _init(args)
    nv, p = args.nves, args.p
    X0 = W.implicit_step_case(nv=nv, p=p)
    Xg = torch.from_numpy(X0).cuda()
    common = dict(dt=1e-3, shear_rate=1.0, bending_modulus=1e-2, tol=1e-6, max_iter=args.max_iter,
                  r_near=args.r_near, solver=args.solver, restart=args.restart)
    nin = 4 * p * (p + 2)
    pc, pc_rec = _train_pc(args, p, common)
    W_, K = args.warmup, args.steps
    timed = dict(common, extrapolate=args.extrapolate)
    # the FP64 operator with the exact far field judges every learned step
    check = I.Stepper(nv, p, precision=0, precond=1, **common)
    learned_kw = dict(timed, precision=args.mixed_precision, precond=2,
                      inner_max_iter=args.inner_max_iter, inner_tol=args.inner_tol,
                      far_coarse=args.far_coarse, far_coarse2=args.far_coarse2,
                      fallback=args.fallback)

    # ---- the two headline forms from the same initial state
    plain, _, first_plain = _run_form(I, nv, p, Xg, dict(timed, precision=0, precond=1), None,
                                      W_ + K, "plain", keep_first=True)
    learned, r64, first_learned = _run_form(I, nv, p, Xg, learned_kw, pc, W_ + K,
                                            "learned+mixed", check=check, keep_first=True)
    plain_t, learned_t = plain[W_:], learned[W_:]
    # ---- attribution (reported, not part of the headline)
    att = {}
    if not args.no_attribution:
        for name, kw, use_pc in (
                ("learned_mixed_exact_far", dict(learned_kw, far_coarse=0.0, far_coarse2=0.0),
                 True),
                ("learned_mixed_two_level_far", dict(learned_kw, far_coarse2=0.0), True),
                ("diagonal_inner_mixed", dict(learned_kw, precond=1), False),
                ("plain_multi_level_far", dict(timed, precision=0, precond=1,
                                               far_coarse=args.far_coarse,
                                               far_coarse2=args.far_coarse2), False)):
            stt, _, _ = _run_form(I, nv, p, Xg, kw, pc if use_pc else None, W_ + K, name)
            stt = stt[W_:]
            att[name] = {"ms_per_step": _mean(stt, "ms_total"),
                         "iterations": [s["iterations"] for s in stt],
                         "inner_iters": [s["inner_iters"] for s in stt],
                         "far_ms_per_step": _mean(stt, "ms_far") + _mean(stt, "ms_far2"),
                         "true_relres": [s["true_relres"] for s in stt]}
    # ---- the headline form with the reference's near zone r_ves 2/sqrt(p_up)
    near_default = None
    if not args.no_attribution:
        stn, r64n, _ = _run_form(I, nv, p, Xg, dict(learned_kw, r_near=0.0), pc, W_ + K,
                                 "learned+mixed r_near default",
                                 check=I.Stepper(nv, p, precision=0, precond=1,
                                                 **dict(common, r_near=0.0)))
        stn = stn[W_:]
        near_default = {"r_near": stn[-1]["r_near"], "near_pairs": [s["near_pairs"] for s in stn],
                        "ms_per_step": _mean(stn, "ms_total"),
                        "near_ms_per_step": _mean(stn, "ms_near"),
                        "near_share": _mean(stn, "ms_near") / _mean(stn, "ms_total"),
                        "iterations": [s["iterations"] for s in stn],
                        "fp64_relres": r64n[W_:]}

    # velocity error of both forms' first step against the FP64 solve of the
    # same system converged to relative residual args.ref_tol (untimed)
    ref = I.Stepper(nv, p, precision=0, precond=1, **dict(common, tol=args.ref_tol))
    ref.set_state(Xg)
    sr = ref.step()
    _log("FP64 reference solve: %d iters, true relres %.2e" % (sr["iterations"], sr["true_relres"]))
    _, Ur, _ = ref.get_state()
    ref.close()
    nr = torch.linalg.norm(Ur)
    vel_err = float(torch.linalg.norm(first_learned[0] - Ur) / nr)
    vel_err_plain = float(torch.linalg.norm(first_plain[0] - Ur) / nr)
    check.close()
    pc_rec["bf16_vs_fp32_on_first_rhs"] = _network_bf16_error(I, nv, p, Xg, learned_kw, pc)

    ms_mixed = _mean(learned_t, "ms_total")
    ms_plain = _mean(plain_t, "ms_total")

    # roofline: the far-field SL direct sum (dominant kernel) of the
    # learned + mixed form; pairs per launch from the step's own count
    pairs = float(learned_t[-1]["far_pairs"])
    flop = pairs * SL_FLOP_PER_PAIR
    far_ms = sum(s["ms_far"] for s in learned_t) / max(1, sum(s["far_calls"] for s in learned_t))
    pairs_plain = float(plain_t[-1]["far_pairs"])
    far_ms_plain = sum(s["ms_far"] for s in plain_t) / max(1, sum(s["far_calls"] for s in plain_t))
    achieved = flop / (far_ms * 1e-3) / 1e12
    scalar = args.far_kernel == "scalar" and args.nves == 512 and args.p == 16
    traffic = FAR32_TRAFFIC.get((args.far_coarse, args.far_coarse2)) if scalar else None
    roof = {"bound": "valu", "achieved": achieved, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
            "frac": achieved / PEAK_FP32_TFLOPS,
            "traffic": traffic[0] if traffic else None,
            "traffic_source": traffic[1] if traffic else None,
            "peak_nonpacked": PEAK_FP32_NONPACKED_TFLOPS,
            "frac_nonpacked": achieved / PEAK_FP32_NONPACKED_TFLOPS,
            "kernel": FAR32_NAMES[args.far_kernel] + " (+k_pack_far32, k_sum_slices)",
            "kernel_ms": far_ms,
            "flop_per_launch": flop, "pairs_per_launch": pairs,
            "two_density_pass_ms": (sum(s["ms_far2"] for s in learned_t)
                                    / max(1, sum(s["far2_calls"] for s in learned_t))),
            "fp64_form": {"kernel": "k_far_f64<SL>" if args.far_kernel == "lds" else "k_far_groups_f64",
                          "kernel_ms": far_ms_plain, "pairs_per_launch": pairs_plain,
                          "achieved": pairs_plain * SL_FLOP_PER_PAIR / (far_ms_plain * 1e-3) / 1e12,
                          "peak": PEAK_FP64_TFLOPS}}

    # the sustained FP32 VALU rate of this GPU, measured in the same run
    # (independent FMA chains): the attainable ceiling beside the published peak
    import ctypes
    r32 = ctypes.c_double()
    ves3d_amd.call("ves3d_hip_valu_probe", 0, ctypes.byref(r32))
    roof["probe_tflops_fp32"] = 2 * r32.value / 1e12
    roof["frac_of_probe"] = achieved / roof["probe_tflops_fp32"]

    cpu = None if args.no_cpu else cpu_baseline_step(nv, p, X0, _mean(plain_t, "matvecs"),
                                                     plain_t[-1]["near_pairs"])
    pc.close()
    return {
        "metric": METRIC if (nv, p) == (512, 16) else METRIC.replace(
            "@512 ves p=16", "@%d ves p=%d" % (nv, p)),
        "value": ms_mixed, "unit": "ms", "n_gpus": 1, "steps": K,
        "warmup": W_, "ms_per_step": ms_mixed, "higher_is_better": False,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32+f64", "data": "synthetic",
        "config": {"workload": "implicit step, %d perturbed ellipsoids (1,1,0.6+-0.05) p=%d, "
                               "20%% volume fraction, shear 1.0, dt 1e-3, %s 1e-6"
                               % (nv, p, {"fgmres": "FGMRES", "bicgstab": "BiCGStab"}[args.solver]),
                   "nves": nv, "p": p, "p_up": 2 * p, "unknowns": nv * nin,
                   "r_near": plain_t[-1]["r_near"]},
        "ms_per_step_plain": ms_plain, "speedup_vs_plain": ms_plain / ms_mixed,
        "learned_form": "precision %d (FP32-compensated far field%s), learned preconditioner in "
                        "inner solves (%d iterations, %g), multi-level far field (p-grid sources "
                        "beyond %g radii, order-p/2 sources beyond %g)" % (
                            args.mixed_precision,
                            ", FP32 singular operator" if args.mixed_precision == 1
                            else ", FP64 singular operator",
                            args.inner_max_iter, args.inner_tol, args.far_coarse,
                            args.far_coarse2),
        "plain_form": "FP64, diagonal spectral preconditioner, upsampled sources for every pair",
        "iterations_learned": [s["iterations"] for s in learned_t],
        "iterations_plain": [s["iterations"] for s in plain_t],
        "initial_guess": ("2 x_{n-1} - x_{n-2} (both forms)" if args.extrapolate
                          else "x_{n-1}, the previous step's solution (both forms)"),
        "inner_solve": {"max_iter": args.inner_max_iter, "tol": args.inner_tol,
                        "inner_iters_learned": [s["inner_iters"] for s in learned_t],
                        "local_matvecs_learned": [s["local_matvecs"] for s in learned_t]},
        "attribution": att,
        "near_default_r_near": near_default,
        "true_relres_learned": [s["true_relres"] for s in learned_t],
        "fp64_relres_learned_every_step": r64,
        "true_relres_plain": [s["true_relres"] for s in plain_t],
        "fell_back": [s["fell_back"] for s in learned_t],
        "fallback": ("diagonal spectral preconditioner (--fallback none: unpreconditioned)"
                     if args.fallback == "diagonal" else "unpreconditioned"),
        "solver": {"outer": args.solver, "restart": args.restart or 30,
                   "true_residual": ("formed from the Arnoldi basis, V (beta e1 - Hbar y)"
                                     if args.solver == "fgmres" else "b - A x, one matvec")},
        "velocity_rel_err_vs_fp64": vel_err, "velocity_rel_err_plain_vs_fp64": vel_err_plain,
        "fp64_reference_relres": sr["true_relres"],
        "network": pc_rec,
        "near_pairs": [s["near_pairs"] for s in plain_t], "r_near": plain_t[-1]["r_near"],
        "stage_ms_plain": {k: _mean(plain_t, k) for k in ("ms_setup", "ms_rhs", "ms_solve",
                                                          "ms_post", "ms_near")},
        "stage_ms_learned": {k: _mean(learned_t, k) for k in ("ms_setup", "ms_rhs", "ms_solve",
                                                              "ms_post", "ms_near")},
        "far_ms_per_step": {"plain": _mean(plain_t, "ms_far") + _mean(plain_t, "ms_far2"),
                            "learned": _mean(learned_t, "ms_far") + _mean(learned_t, "ms_far2")},
        "roofline": roof, "cpu_baseline": cpu,
    }


def bench_trajectory(args):
    """The step workload run for --traj-steps steps in both forms from the
    same initial state (the multi-step form of the full-step configuration):
    per-step ms, outer and inner iterations, fallbacks and true residuals of
    every step, and how far the two trajectories' positions drift apart."""
    import torch
    import ves3d_amd
    from ves3d_amd import implicit as I
    from ves3d_amd import workloads as W

    _init(args)
    nv, p, nsteps = args.nves, args.p, args.traj_steps
    X0 = W.implicit_step_case(nv=nv, p=p)
    Xg = torch.from_numpy(X0).cuda()
    common = dict(dt=1e-3, shear_rate=1.0, bending_modulus=1e-2, tol=1e-6, max_iter=args.max_iter,
                  r_near=args.r_near, solver=args.solver, restart=args.restart)
    pc = _train_pc(args, p, common)[0]
    forms = {
        "plain": I.Stepper(nv, p, precision=0, precond=1, **common),
        "learned": I.Stepper(nv, p, precision=args.mixed_precision, precond=2,
                             inner_max_iter=args.inner_max_iter, inner_tol=args.inner_tol,
                             far_coarse=args.far_coarse, far_coarse2=args.far_coarse2,
                             **common),
    }
    I.attach(forms["learned"], pc, collect=False)
    out, final = {}, {}
    for name, st in forms.items():
        st.set_state(Xg)
        rec = []
        for k in range(nsteps):
            s = st.step()
            rec.append(s)
            _log("%s step %d: %.1f ms, %d iters (%d inner), true %.2e fell_back %d" %
                 (name, k, s["ms_total"], s["iterations"], s["inner_iters"], s["true_relres"],
                  s["fell_back"]))
        final[name] = st.get_state()[0].clone()
        warm = rec[1:] if len(rec) > 1 else rec
        out[name] = {
            "ms_per_step_warm_mean": float(np.mean([s["ms_total"] for s in warm])),
            "ms_per_step_warm_max": float(np.max([s["ms_total"] for s in warm])),
            "iterations": [s["iterations"] for s in rec],
            "inner_iters": [s["inner_iters"] for s in rec],
            "fell_back_steps": int(sum(s["fell_back"] for s in rec)),
            "max_true_relres": float(np.max([s["true_relres"] for s in rec])),
            "all_converged": bool(all(s["status"] == 100 for s in rec)),
        }
        st.close()
    pc.close()
    disp = float(torch.linalg.norm(final["plain"] - Xg))
    drift = float(torch.linalg.norm(final["learned"] - final["plain"]))
    ms_l = out["learned"]["ms_per_step_warm_mean"]
    ms_p = out["plain"]["ms_per_step_warm_mean"]
    return {
        "metric": "ms per implicit step over a %d-step trajectory @%d ves p=%d "
                  "(learned-PC+mixed vs plain) at 1e-6" % (nsteps, nv, p),
        "value": ms_l, "unit": "ms", "n_gpus": 1, "steps": nsteps, "higher_is_better": False,
        "data": "synthetic",
        "config": {"workload": "implicit-step trajectory", "nves": nv, "p": p, "steps": nsteps},
        "ms_per_step_plain": ms_p, "speedup_vs_plain": ms_p / ms_l,
        "position_drift_vs_plain": drift / max(disp, 1e-300),
        "position_displacement_plain": disp, "forms": out,
    }


def _near_cost_per_pair(p):
    """Oracle seconds per near-singular pair: setup (search, projection) and
    one application (interpolation minus direct sum), measured on two
    perturbed spheres 0.05 apart."""
    from oracle import near as N
    from oracle import ves
    from ves3d_amd import workloads as W
    X = W.perturbed_shapes(2, p, amp=0.05, seed=11)
    X[1] += np.array([2.05, 0.0, 0.0])[:, None]
    st = ves.StokesSurface(X, p, 2 * p, near=False)
    ng = 2 * p * (p + 1)
    t0 = time.perf_counter()
    nr = N.NearSingular(st.scoord_far, st.trg, 2, 2 * p, ng)
    t_setup = time.perf_counter() - t0
    st.near = nr
    _, V, q, _, _, _ = st.terms(np.random.default_rng(1).uniform(-1, 1, X.shape))
    t0 = time.perf_counter()
    st.near_velocity(q, V)
    t_apply = time.perf_counter() - t0
    n = max(1, len(nr.pairs))
    return t_setup / n, t_apply / n


def cpu_baseline_step(nv, p, X, n_matvec, near_pairs):
    """The FP64 oracle on the host, on a bounded sample, extrapolated to one
    step of the same workload: per-vesicle singular-operator build (2
    vesicles timed) x nv, plus n_matvec x (far field: all sources against a
    sample of 256 targets, scaled to all targets; local matvec work: one
    oracle matvec on 2 vesicles, scaled by nv/2), plus the near-singular
    correction per pair (two close vesicles timed) x the step's pair count."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as OL
    from oracle import sph, ves
    t0 = time.perf_counter()
    sph.stokes_singular_integ(X[:2], p, 2 * p, want_dl=False)
    t_sing = (time.perf_counter() - t0) / 2
    st = ves.StokesSurface(X[:2], p, 2 * p)  # small: the far geometry of 2 vesicles
    # far field sample at full size
    pu = 2 * p
    mu = 2 * pu * (pu + 1)
    shc = sph.grid2shc(X.reshape(-1, 2 * p * (p + 1)), p, p)
    up = sph.shc2grid(shc, p, pu).reshape(nv, 3, mu)
    src = np.ascontiguousarray(np.transpose(up, (0, 2, 1)).reshape(-1, 3))
    den = np.ones_like(src) * 1e-3
    ng = 2 * p * (p + 1)
    pairs_matvec = float(nv * ng) * (nv - 1) * (mu + 2)  # far pairs per matvec, exact far field
    ref = OL.load_ref()
    if ref is not None:
        # the far-field leg on the reference's own StokesAlltoAll
        # (src/CPUKernels.cc, compiled from source into oracle/_ref): all-to-all
        # among the first n_all upsampled sources, scaled by pairs
        n_all = min(src.shape[0], 60000)
        s_all = np.ascontiguousarray(src[:n_all])
        d_all = np.ascontiguousarray(den[:n_all])
        out = np.zeros((n_all, 3))
        t0 = time.perf_counter()
        ref.ref_stokes_all2all_f64(OL.p(s_all), OL.p(d_all), OL.cs(n_all), OL.p(out))
        t_far = (time.perf_counter() - t0) * pairs_matvec / (float(n_all) * n_all)
        far_leg = ("the reference's own StokesAlltoAll (src/CPUKernels.cc built from source "
                   "into oracle/_ref) all-to-all on %d upsampled sources, scaled by pairs to "
                   "%.3g pairs per matvec" % (n_all, pairs_matvec))
    else:
        ntr = 256
        trg = np.ascontiguousarray(np.transpose(X, (0, 2, 1)).reshape(-1, 3)[:: (nv * ng) // ntr][:ntr])
        out = np.zeros((ntr, 3))
        orc = OL.load()
        t0 = time.perf_counter()
        orc.orc_far_field(OL.p(src), OL.p(den), None, OL.cs(src.shape[0]), OL.p(trg), OL.p(out),
                          OL.cs(ntr), OL.cs(0), OL.cs(0))
        t_far = (time.perf_counter() - t0) * (nv * ng) / ntr
        far_leg = ("the oracle's C far field of all %d sources on 256 targets (x%d)" %
                   (src.shape[0], (nv * ng) // ntr))
    imp = ves.Implicit(X[:2], p, dt=1e-3, shear=1.0, near=False)
    x = np.random.default_rng(0).standard_normal(4 * 2 * p * (p + 2))
    t0 = time.perf_counter()
    imp.matvec(x)
    t_local = (time.perf_counter() - t0) * nv / 2
    near_setup, near_apply = _near_cost_per_pair(p)
    total = t_sing * nv + n_matvec * (t_far + t_local) + \
        near_pairs * (near_setup + n_matvec * near_apply)
    del st
    far_share = n_matvec * t_far / total
    return {"value": total * 1e3, "unit": "ms per implicit step (extrapolated)",
            "cores": _cpu_threads(), "kind": "reference" if ref is not None else "port",
            "sample": "far field (%.0f%% of the total): %s; the rest on the FP64 oracle port "
                      "(numpy): singular operator of 2 vesicles (x%d), one 2-vesicle matvec "
                      "(x%d), near-singular setup + %.0f applications per pair on two close "
                      "vesicles (x%d pairs); all times %.0f matvecs per step (the GPU plain "
                      "run's count)" % (100 * far_share, far_leg, nv, nv // 2, n_matvec,
                                        near_pairs, n_matvec),
            "components_s": {"singular_per_vesicle": t_sing, "far_per_matvec": t_far,
                             "local_per_matvec": t_local, "near_setup_per_pair": near_setup,
                             "near_apply_per_pair": near_apply}}


# ---------------------------------------------------------------- far only
def bench_matvec(args):
    """One prepared stepper of the step workload and K Galerkin matvecs (each
    one far-field launch between vesicles plus the local operator), after a
    device copy of known size whose HBM counters calibrate the far kernel's
    (FETCH_SIZE / WRITE_SIZE under rocprofv3 --pmc)."""
    import torch
    import ves3d_amd
    from ves3d_amd import implicit as I
    from ves3d_amd import workloads as W

    _init(args)
    nv, p = args.nves, args.p
    prec = args.mixed_precision if args.dtype == "f32c" else 0
    # the mixed form's kernel with its multi-level far field, the FP64 form's
    # over the upsampled sources (as the step workload's two headline forms)
    coarse = args.far_coarse if prec else 0.0
    coarse2 = args.far_coarse2 if prec else 0.0
    st = I.Stepper(nv, p, dt=1e-3, shear_rate=1.0, precision=prec, precond=1, r_near=args.r_near,
                   far_coarse=coarse, far_coarse2=coarse2)
    st.set_state(torch.from_numpy(W.implicit_step_case(nv=nv, p=p)).cuda())
    st.prepare()
    pairs = st.far_pairs()
    x = torch.from_numpy(np.random.default_rng(3).standard_normal(st.n)).cuda()
    cal_bytes = 1 << 29                          # 512 MiB read + 512 MiB written
    a = torch.ones(cal_bytes // 8, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    b.copy_(a)
    torch.cuda.synchronize()
    _log("calibration copy: %d bytes read, %d bytes written" % (cal_bytes, cal_bytes))
    for _ in range(args.warmup):
        st.matvec(x)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        st.matvec(x)
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) / max(1, args.steps) * 1e3
    st.close()
    return {
        "metric": "ms per Galerkin matvec @512 ves p=16", "value": ms, "unit": "ms", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": False,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32+f64" if prec else "f64",
        "data": "synthetic",
        "config": {"workload": "implicit-step matvec, 512 perturbed ellipsoids p=16, 20% volume "
                               "fraction", "precision": prec, "far_kernel": args.far_kernel,
                   "far_coarse": coarse, "far_coarse2": coarse2, "far_pairs": pairs},
        "roofline": None, "cpu_baseline": None,
    }


# singular build, SL + DL, per vesicle: 2 M N K summed over the GEMMs of
# the reference's dense-DFT algorithm (RotateAll, upsampling SHT, pole
# kernel, transpose SHT, RotateTranspose, final Grid2SHC)
SING_GFLOP_PER_VES = {16: 1.772, 24: 12.74}
PEAK_HBM_TBS = 8.0


def _timed(fn, reps=3):
    import torch
    fn()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    best = 1e30
    for _ in range(reps):
        e0.record()
        fn()
        e1.record()
        torch.cuda.synchronize()
        best = min(best, e0.elapsed_time(e1))
    return best


def bench_stages(args):
    """Per-stage rooflines outside the far field: the batched singular build
    (SL + DL; 2048 vesicles at p = 16 and 512 at p = 24, FP64 and FP32
    transforms, with the FP32 form's error per output), the self-apply GEMV
    and the SH transforms of the 512-vesicle step."""
    import torch
    import ves3d_amd
    from ves3d_amd import sph as S
    from ves3d_amd import workloads as W
    _init(args)
    out = {}
    rng = np.random.default_rng(11)
    for nv, p in ((2048, 16), (512, 24)):
        X = torch.from_numpy(W.perturbed_shapes(nv, p, amp=0.1, seed=21)).cuda()
        n = 3 * S.ncoef(p)
        F = torch.from_numpy(rng.uniform(-1, 1, (nv, n))).cuda()   # traction coefficients
        res = {}
        mats = {}
        for dt, name in ((torch.float64, "f64"), (torch.float32, "f32")):
            SL = torch.empty((nv, n, n), dtype=dt, device="cuda")
            DL = torch.empty((nv, n, n), dtype=dt, device="cuda")
            ms = _timed(lambda: S.singular_build(X, nv, p, 2 * p, SL, DL), reps=2)
            gf = SING_GFLOP_PER_VES[p] * nv
            peak = PEAK_FP64_TFLOPS if dt == torch.float64 else PEAK_FP32_TFLOPS
            res[name] = {"ms": ms, "tflops": gf / ms, "frac_of_peak": gf / ms / peak, "peak": peak}
            V = torch.empty((nv, n), dtype=torch.float64, device="cuda")
            Vd = torch.empty_like(V)
            S.self_apply(SL, F, None, None, nv, p, V)
            S.self_apply(None, None, DL, F, nv, p, Vd)
            torch.cuda.synchronize()
            mats[name] = (SL, DL, V, Vd)
            _log("singular build %d ves p=%d %s: %.1f ms, %.1f TFLOP/s" % (nv, p, name, ms, gf / ms))
        SLa, DLa, Va, Vda = mats["f64"]
        SLb, DLb, Vb, Vdb = mats["f32"]
        rel = lambda a, b: float((a.double() - b).abs().max() / b.abs().max())
        res["f32_error_vs_f64"] = {"SL_matrix": rel(SLb, SLa), "DL_matrix": rel(DLb, DLa),
                                   "SL_velocity": rel(Vb, Va), "DL_velocity": rel(Vdb, Vda)}
        out["singular_build_%dves_p%d" % (nv, p)] = res
        if nv == 2048:
            # the self-apply GEMV of 512 vesicles at p = 16 in FP64 (HBM-bound)
            SL5 = SLa[:512].contiguous()
            F5, V5 = F[:512].contiguous(), torch.empty((512, n), dtype=torch.float64, device="cuda")
            ms = _timed(lambda: S.self_apply(SL5, F5, None, None, 512, p, V5), reps=5)
            byts = SL5.numel() * 8 + 2 * F5.numel() * 8
            out["self_apply_512ves_p16_f64"] = {"ms": ms, "TBps": byts / ms / 1e9,
                                                "frac_of_hbm": byts / ms / 1e9 / PEAK_HBM_TBS,
                                                "bytes": byts}
            del SL5
        del mats, SLa, DLa, SLb, DLb
        torch.cuda.empty_cache()
    # SH transforms of the step: 3 x 512 functions, p = 16 grid -> coefficients,
    # coefficients -> p = 32 grid (the far-field density upsampling)
    p, N = 16, 3 * 512
    g = torch.from_numpy(rng.standard_normal((N, S.ngrid(p)))).cuda()
    c = torch.empty((N, S.ncoef(p)), dtype=torch.float64, device="cuda")
    u = torch.empty((N, S.ngrid(2 * p)), dtype=torch.float64, device="cuda")
    ms = _timed(lambda: S.grid2shc(g, N, p, p, c), reps=5)
    byts = (g.numel() + c.numel()) * 8
    out["grid2shc_1536f_p16"] = {"ms": ms, "TBps": byts / ms / 1e9, "bytes": byts}
    ms = _timed(lambda: S.shc2grid(c, N, p, 2 * p, u), reps=5)
    byts = (c.numel() + u.numel()) * 8
    out["shc2grid_1536f_p16_to_p32"] = {"ms": ms, "TBps": byts / ms / 1e9, "bytes": byts}
    return {"metric": "per-stage rooflines (singular build, self apply, SH transforms)",
            "value": out["singular_build_2048ves_p16"]["f64"]["ms"], "unit": "ms",
            "n_gpus": 1, "higher_is_better": False, "data": "synthetic",
            "config": {"workload": "stages"}, "stages": out}


def bench_learned(args):
    """The headline form alone (learned preconditioner + mixed precision +
    multi-level far field) for warmup + steps steps: a short program for
    kernel traces and counters of the form that the step workload times."""
    import torch
    import ves3d_amd
    from ves3d_amd import implicit as I
    from ves3d_amd import workloads as W
    _init(args)
    nv, p = args.nves, args.p
    Xg = torch.from_numpy(W.implicit_step_case(nv=nv, p=p)).cuda()
    common = dict(dt=1e-3, shear_rate=1.0, bending_modulus=1e-2, tol=1e-6, max_iter=args.max_iter,
                  r_near=args.r_near, solver=args.solver, restart=args.restart)
    pc, pc_rec = _train_pc(args, p, common)
    kw = dict(common, extrapolate=args.extrapolate, precision=args.mixed_precision, precond=2,
              inner_max_iter=args.inner_max_iter, inner_tol=args.inner_tol,
              far_coarse=args.far_coarse, far_coarse2=args.far_coarse2)
    stats, _, _ = _run_form(I, nv, p, Xg, kw, pc, args.warmup + args.steps, "learned+mixed")
    st = stats[args.warmup:]
    ms = _mean(st, "ms_total")
    return {"metric": "ms per implicit step @512 ves p=16, learned-PC+mixed form alone",
            "value": ms, "unit": "ms", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
            "ms_per_step": ms, "higher_is_better": False, "data": "synthetic",
            "config": {"workload": "learned form alone", "nves": nv, "p": p},
            "iterations": [s["iterations"] for s in st], "network": pc_rec,
            "stage_ms": {k: _mean(st, k) for k in ("ms_setup", "ms_rhs", "ms_solve", "ms_post",
                                                   "ms_near", "ms_far", "ms_far2")},
            "far_pairs": st[-1]["far_pairs"]}


def bench_build(args):
    """One singular-operator build (SL + DL) of args.nves vesicles at args.p in
    FP64 and in FP32 transforms: a short program for hardware-counter runs
    (MFMA busy cycles and MFMA flops of the build's GEMMs)."""
    import torch
    import ves3d_amd
    from ves3d_amd import sph as S
    from ves3d_amd import workloads as W
    _init(args)
    nv, p = args.nves, args.p
    X = torch.from_numpy(W.perturbed_shapes(nv, p, amp=0.1, seed=21)).cuda()
    n = 3 * S.ncoef(p)
    out = {}
    for dt, name in ((torch.float64, "f64"), (torch.float32, "f32")):
        SL = torch.empty((nv, n, n), dtype=dt, device="cuda")
        DL = torch.empty((nv, n, n), dtype=dt, device="cuda")
        ms = _timed(lambda: S.singular_build(X, nv, p, 2 * p, SL, DL), reps=1)
        gf = SING_GFLOP_PER_VES.get(p, float("nan")) * nv
        out[name] = {"ms": ms, "tflops_algorithmic": gf / ms}
        del SL, DL
        torch.cuda.empty_cache()
    return {"metric": "singular build (SL + DL) ms", "value": out["f64"]["ms"], "unit": "ms",
            "n_gpus": 1, "higher_is_better": False, "data": "synthetic",
            "config": {"workload": "singular build", "nves": nv, "p": p}, "build": out}


def bench_probe(args):
    """The VALU rate this GPU sustains (independent FMA chains, FP32 non-packed
    and FP64): the attainable ceiling beside the published peaks."""
    import ctypes
    import ves3d_amd
    _init(args)
    r32, r64 = ctypes.c_double(), ctypes.c_double()
    ves3d_amd.call("ves3d_hip_valu_probe", 0, ctypes.byref(r32))
    ves3d_amd.call("ves3d_hip_valu_probe", 1, ctypes.byref(r64))
    return {"metric": "sustained lane FMA/s (FP32 non-packed, FP64)", "value": r32.value,
            "unit": "lane FMA/s", "n_gpus": 1, "fp64_lane_fma_per_s": r64.value,
            "fp32_tflops": 2 * r32.value / 1e12, "fp64_tflops": 2 * r64.value / 1e12,
            "higher_is_better": True, "data": "synthetic", "config": {"workload": "probe"}}


def bench_far(args):
    import torch
    import ves3d_amd
    from ves3d_amd import stokes as K
    from ves3d_amd import workloads as W

    _init(args)
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
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="step",
                    choices=["step", "far", "matvec", "probe", "stages", "build", "trajectory",
                             "learned"])
    ap.add_argument("--dtype", default="f64", choices=["f64", "f32c"])
    ap.add_argument("--nves", type=int, default=512)
    ap.add_argument("--p", type=int, default=16)
    ap.add_argument("--max-iter", type=int, default=300)
    ap.add_argument("--width", type=int, default=256)
    ap.add_argument("--capacity", type=int, default=200000)
    ap.add_argument("--epochs", type=int, default=20)
    ap.add_argument("--train-nves", type=int, default=256)
    ap.add_argument("--train-steps", type=int, default=2,
                    help="training suspensions (one plain step each) with --pc-weights train")
    ap.add_argument("--pc-weights", default="stored",
                    help="'stored' (the committed network), a path without suffix, or 'train'")
    ap.add_argument("--activations", default="bf16", choices=["fp32", "bf16"],
                    help="the network's activations (BF16 on BF16 MFMA, or FP32)")
    ap.add_argument("--ref-tol", type=float, default=1e-10)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--inner-max-iter", type=int, default=6,
                    help="learned form: inner BiCGStab iterations on the vesicles' own operators "
                         "per preconditioner application (0: apply the network once)")
    ap.add_argument("--inner-tol", type=float, default=1e-2)
    ap.add_argument("--traj-steps", type=int, default=100,
                    help="trajectory workload: steps per form")
    ap.add_argument("--extrapolate", action="store_true",
                    help="every timed form starts each solve from 2 x_{n-1} - x_{n-2} instead of "
                         "x_{n-1} (the reference's initial guess); needs --warmup >= 2 for "
                         "steady-state steps")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--mixed-precision", type=int, default=2, choices=[1, 2],
                    help="precision of the learned form: 1 FP32 singular operator + FP32-"
                         "compensated far field, 2 FP64 singular operator + FP32-compensated far field")
    ap.add_argument("--far-kernel", default="scalar", choices=sorted(FAR_KINDS),
                    help="far field between vesicles: sources through the scalar cache (scalar2: "
                         "two targets per lane in FP32) or LDS tiles")
    ap.add_argument("--far-coarse", type=float, default=3.0,
                    help="learned form: two-level far field, p-grid sources for waves of "
                         "targets beyond this many vesicle radii (0: upsampled sources "
                         "everywhere, as the plain form)")
    ap.add_argument("--fallback", default="diagonal", choices=["diagonal", "none"],
                    help="learned form: the solve that repeats a failed learned solve, "
                         "diagonal-preconditioned or unpreconditioned")
    ap.add_argument("--far-coarse2", type=float, default=6.0,
                    help="learned form: third far-field level, order-p/2 sources for waves "
                         "beyond this many radii (0: two levels)")
    ap.add_argument("--no-attribution", action="store_true",
                    help="step workload: skip the attribution forms and the near-default line")
    ap.add_argument("--pole-kernel", default="fused", choices=["fused", "separate"],
                    help="the FP64 singular build's pole kernel: fused with its transposed "
                         "synthesis, or the separate kernel and grouped GEMMs")
    ap.add_argument("--solver", default="fgmres", choices=["bicgstab", "fgmres"],
                    help="outer solver of every form: flexible GMRES (the reference's implicit "
                         "solve is PETSc GMRES) or BiCGStab with a true-residual check")
    ap.add_argument("--restart", type=int, default=0, help="FGMRES restart length (0: 30)")
    ap.add_argument("--r-near", type=float, default=0.1,
                    help="near-singular zone radius (0: the reference's r_ves 2/sqrt(p_up))")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: only the one-GPU form is built; --gpus %d is not supported "
              "(across GPUs the path would shard by vesicle)" % args.gpus, file=sys.stderr)
        sys.exit(2)
    res = {"step": bench_step, "far": bench_far, "matvec": bench_matvec,
           "probe": bench_probe, "stages": bench_stages, "build": bench_build,
           "trajectory": bench_trajectory, "learned": bench_learned}[args.workload](args)
    print(json.dumps(res))


if __name__ == "__main__":
    main()

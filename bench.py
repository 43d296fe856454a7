"""Benchmark: time to relative data residual 1e-3 of helical OS-SART at 512x512x256,
learned warm start + BF16 storage vs plain FP32, on one GPU.

One step is one OS-SART sweep (20 subsets; every view forward- and back-projected
once) over the clinical-size helical scan: 512x512x256 volume at 0.7 mm, 888x64
detector, 984 views/turn x 4 turns (3936 views), pitch 0.9, FOV mask r^2 < 0.99.
The measured sinogram b = A x_true (FP32, noise-free) of a seeded 40-ellipsoid
phantom is resident in HBM before timing starts; x_0 = 0.

Two forms are timed on the same b, each from x_0 = 0 until the relative data residual
||b - A x|| / ||b|| (FP64 sums, accumulated during each sweep) first reaches the
target, with the row / column-sum set-up inside the timed region:
  * plain FP32 OS-SART;
  * learned + BF16: BF16 projector operands (FP32 master iterate), two sweeps, the
    warm-start network's predicted update, then plain OS-SART sweeps (with the
    fallback to plain OS-SART when the residual does not drop).
``value`` is the learned + BF16 time to the target (ctrecon_amd/weights/warm_start.npz;
the later-stage network warm_start_stage2.npz acts, when present, --after2 sweeps after
the first network's update).  Without trained weights ``value`` is null and the plain
time is reported alone.

Steps: --warmup W untimed plain FP32 sweeps, then the plain FP32 solve from x_0 = 0 whose
first K = --steps sweeps are the timed steps (``steps`` = K, ``ms_per_step`` = their mean
GPU time; if the target is met in fewer sweeps, the solve continues to K sweeps for these
statistics, outside the time to target).  Both forms always run to the target (at most
--max-steps sweeps).

Prints ONE JSON line on stdout; progress goes to stderr.

    python bench.py [--steps K] [--warmup W] [--target 1e-3] [--form both|plain|learned] [--no-cpu]
    python bench.py --sweeps-only --steps K    # exactly K plain sweeps (profiling runs)
"""

from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = ("time to rel. residual 1e-3, 512²×256 helical OS-SART: learned+BF16 vs "
          "plain FP32")
HBM_PEAK = 8.0e12  # MI355X, bytes/s


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_model() -> str:
    """The host CPU's model name (as lscpu prints it), from /proc/cpuinfo."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(g, vol, n_views, target_sweeps, setup_sweeps=1.0):
    """The oracle's FP32 restatement of the reference path (integral images + bilinear
    samples, OpenMP over the host cores) on a bounded sample of the same scan."""
    from oracle import oracle as O

    views = np.linspace(0, g.PN - 1, n_views).round().astype(np.int64)
    gs = g.subset(views)
    t0 = time.perf_counter()
    prj = O.fp_f32(gs, vol)
    t1 = time.perf_counter()
    O.bp_f32(gs, prj)
    t2 = time.perf_counter()
    per_view = (t2 - t0) / n_views  # one forward + one back projection of one view
    out = {
        "unit": "s",
        "cores": O.num_threads(),
        "cpu_model": cpu_model(),
        "logical_cpus": os.cpu_count(),
        "kind": "port",
        "projections_per_s": 1.0 / per_view,
        "fp_s": t1 - t0,
        "bp_s": t2 - t1,
        "sample": (f"{n_views} views spread over the 3936-view scan, FP32 forward + back "
                   f"projection of the full 512x512x256 volume ({t2 - t0:.1f} s of CPU work); "
                   f"value = (plain sweeps to target + 1 set-up sweep) x 3936 views x "
                   f"measured s/view, extrapolated"),
        "target_sweeps": target_sweeps,
    }
    out["value"] = (None if target_sweeps is None else
                    (target_sweeps + setup_sweeps) * g.PN * per_view)
    return out


def headline_record(args, plain, learned, learned_sino, roof, pps, kernels, cpu, have_net,
                    views, recipe=None):
    """The one-line JSON record of the clinical workload.  `plain` / `learned` are the legs'
    result dicts (learned None without trained weights); `plain["timed_ms"]` holds the K
    timed sweeps."""
    timed = plain["timed_ms"]
    t_plain = plain["time_to_target_s"]
    t_learned = learned["time_to_target_s"] if learned is not None else None
    both = t_plain is not None and t_learned is not None
    form = ("learned warm start + BF16 storage (FP32 accumulation)" if learned is not None else
            "plain FP32" + ("" if have_net else " (no trained network weights found)"))
    out = {
        "metric": METRIC,
        "value": t_learned,
        "unit": "s",
        "n_gpus": 1,
        "steps": len(timed),
        "warmup": args.warmup,
        "ms_per_step": float(np.mean(timed)) if len(timed) else None,
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 accumulation, bf16 storage" if learned is not None else "f32",
        "data": "synthetic: seeded 40-ellipsoid phantom (seed 1) + uniform texture; "
                "b = A x_true (noise-free FP32)",
        "config": {
            "workload": "512x512x256 helical OS-SART to rel. residual "
                        f"{args.target:g}, " + ("reference back projector" if args.bp == "ref"
                                                 else "matched adjoint"),
            "form": form,
            "step": "one plain FP32 OS-SART sweep (20 subsets, 3936 views forward- and "
                    "back-projected)",
            "volume": [512, 512, 256], "voxel_mm": 0.7, "detector": [888, 64],
            "views": int(views), "views_per_turn": 984, "pitch": 0.9,
            "subsets": args.subsets, "target_rel_residual": args.target,
        },
        "roofline": roof,
        "sweeps_to_target": learned["sweeps"] if learned is not None else plain[
            "sweeps_to_target"],
        "plain_fp32": {k: v for k, v in plain.items() if k != "timed_ms"},
        "learned_bf16": learned,
        "learned_bf16_sinogram_only": learned_sino,
        # the configured recipe's network beside the shipped ones (not the reported value)
        "learned_bf16_configured_recipe": recipe,
        "speedup_vs_plain": t_plain / t_learned if both else None,
        "rmse_ratio_vs_plain": (learned["rmse_vs_truth_in_mask"] / plain["rmse_vs_truth_in_mask"]
                                if learned is not None else None),
        "projections_per_s": pps,
        "kernel_avg_ms": kernels,
    }
    if cpu is not None:
        out["cpu_baseline"] = cpu
    return out


def batched_workload(args):
    """Config-4 form: 16 helical scans of uneven size, seeds 100-115 (volumes
    U{256..512}^2 x U{64..192}, voxels U(0.5, 1.0) mm, pitch U(0.5, 1.5), turns U{2,3,4};
    888x64 detector, 984 views/turn), BF16 storage, 20 subsets.  Times one OS-SART sweep
    of the whole batch (one launch per subset step and kernel, HIP-graph replay, tuned
    launch shapes) against the same sweeps run scan by scan."""
    from ctrecon_amd import geometry as G
    from ctrecon_amd import phantom as P
    from ctrecon_amd.batch import Batch
    from ctrecon_amd.plan import DeviceArray, Plan, Sart

    plans, xs, bs, desc = [], [], [], []
    views = 0
    for seed in range(100, 116):
        rng = np.random.default_rng(seed)
        nxy = int(rng.integers(256, 513)) // 4 * 4
        nz = int(rng.integers(64, 193)) // 4 * 4
        vox = float(rng.uniform(0.5, 1.0))
        pitch = float(rng.uniform(0.5, 1.5))
        turns = int(rng.integers(2, 5))
        cfg = G.helical_config(sid=541, sdd=949, DNU=888, DNV=64, col_width=1.0239,
                               col_height=1.096439, views_per_turn=984, turns=turns,
                               feed_mm_per_turn=pitch * 64 * 0.625, N_xy=nxy, N_z=nz,
                               voxel_xy=vox, voxel_z=vox, col_height_at_iso=0.625)
        g = G.dd3_geometry(cfg, mask=G.fov_mask(nxy))
        vol = P.ellipsoid_mixture(g, seed=seed, n=12)
        plan = Plan(g, 0)
        plan.set_storage("bf16")
        b = DeviceArray(g.prj_shape)
        plan.fp(DeviceArray.from_host(vol), b)
        x = DeviceArray(g.vol_shape)
        x.zero()
        plans.append(plan)
        xs.append(x)
        bs.append(b)
        views += g.PN
        desc.append([nxy, nz, round(vox, 3), round(pitch, 3), turns])
        log(f"scan {seed}: {nxy}x{nxy}x{nz} @ {vox:.3f} mm, pitch {pitch:.2f}, {turns} turns")
    K = max(args.steps, 3)
    # scan by scan (sessions with the same storage)
    seq = 0.0
    for plan, x, b in zip(plans, xs, bs):
        S = Sart(plan, b, args.subsets)
        S.sweeps(x, 1)
        _, ms, _ = S.sweeps(x, K)
        seq += float(np.mean(ms))
        S.close()
    for x in xs:
        x.zero()
    t0 = time.perf_counter()
    B = Batch(plans, xs, bs, args.subsets)
    create_s = time.perf_counter() - t0
    B.sweeps(1)  # graph capture + first replay
    res, ms = B.sweeps(K)
    graph_ms = float(np.mean(ms))
    B.set_graph(False)
    _, ms2 = B.sweeps(K)
    info = B.info()
    B.close()
    out = {
        "metric": "batched OS-SART sweep, 16 helical scans of uneven size (BF16 storage)",
        "value": graph_ms,
        "unit": "ms per sweep",
        "n_gpus": 1,
        "steps": K,
        "higher_is_better": False,
        "dtype": "f32 accumulation, bf16 storage",
        "data": "synthetic: seeded 12-ellipsoid phantoms, b = A x_true",
        "config": {"workload": "16 scans, seeds 100-115, 888x64 detector, 984 views/turn, "
                               "20 subsets", "scans": desc, "views_total": views},
        "sequential_ms_per_sweep": seq,
        "launches_ms_per_sweep": float(np.mean(ms2)),
        "speedup_vs_sequential": seq / graph_ms,
        "projections_per_s": views / (graph_ms / 1e3),
        "tuning": info,
        "create_s": create_s,
        "residual_after_sweeps": res[-1].tolist(),
    }
    print(json.dumps(out))


def projectors_workload(args):
    """Every projector model of the library on the clinical geometry (512x512x256, 888x64,
    3936 views): one OS-SART subset (every 20th view, 197 views) forward- and back-projected
    per launch.  Times each launch (mean of K after one warm-up, synchronised wall clock
    around the launches) and prices it with the byte model B_view = 4 N_cone + 4 DNU DNV
    (N_cone: masked voxels inside the view's cone)."""
    from ctrecon_amd import geometry as G
    from ctrecon_amd.plan import DeviceArray, Plan, to_bf16

    g = G.dd3_geometry(G.config_clinical(), mask=G.fov_mask(512))
    plan = Plan(g, 0)
    rng = np.random.default_rng(0)
    vs = args.subsets
    vc = len(range(0, g.PN, vs))
    dv = DeviceArray.from_host(rng.uniform(0, 1, g.vol_shape).astype(np.float32))
    dp = DeviceArray.from_host(rng.uniform(0, 1, (vc, g.DNU, g.DNV)).astype(np.float32))
    dvo = DeviceArray(g.vol_shape)
    dv16, dp16 = to_bf16(dv), to_bf16(dp)
    cone = float(np.sum(plan.cone_voxels(0, vs, vc)))
    bytes_launch = 4.0 * cone + 4.0 * g.DNU * g.DNV * vc
    K = max(args.steps, 3)
    legs = [
        ("fp_branchless_direct", "direct", lambda: plan.fp(dv, dp, 0, vs, vc)),
        ("bp_branchless_reference", "direct", lambda: plan.bp(dp, dvo, 0, vs, vc, "ref")),
        ("bp_matched_adjoint", "direct", lambda: plan.bp(dp, dvo, 0, vs, vc, "matched")),
        ("fp_branchless_bf16", "direct", lambda: plan.fp_bf16(dv16, dp16, 0, vs, vc)),
        ("bp_branchless_bf16", "direct", lambda: plan.bp_bf16(dp16, dvo, 0, vs, vc, "ref")),
        ("fp_branchless_sat", "sat", lambda: plan.fp(dv, dp, 0, vs, vc)),
        ("bp_branchless_sat", "sat", lambda: plan.bp(dp, dvo, 0, vs, vc, "ref")),
        ("fp_pseudo_dd", "pseudo_dd", lambda: plan.fp(dv, dp, 0, vs, vc)),
        ("bp_pseudo_dd", "pseudo_dd", lambda: plan.bp(dp, dvo, 0, vs, vc, "ref")),
        ("bp_zline", "zline", lambda: plan.bp(dp, dvo, 0, vs, vc, "ref")),
    ]
    table = {}
    for name, method, fn in legs:
        plan.set_method(method)
        fn()
        plan.synchronize()
        t0 = time.perf_counter()
        for _ in range(K):
            fn()
        plan.synchronize()
        ms = (time.perf_counter() - t0) / K * 1e3
        table[name] = {"ms": round(ms, 3), "projections_per_s": round(vc / (ms / 1e3), 1),
                       "hbm_frac": round(bytes_launch / (ms / 1e3) / HBM_PEAK, 4)}
        log(f"{name}: {ms:.2f} ms per {vc}-view launch")
    out = {
        "metric": "projections per second per projector model, 512x512x256 helical, "
                  "888x64 detector",
        "value": table["fp_branchless_direct"]["projections_per_s"],
        "unit": "projections/s",
        "n_gpus": 1,
        "steps": K,
        "higher_is_better": True,
        "dtype": "f32",
        "data": "synthetic: uniform random volume and sinogram",
        "config": {"workload": f"one {vc}-view subset (every {vs}th view of 3936) per launch"},
        "bytes_per_launch_model": bytes_launch,
        "kernels": table,
    }
    print(json.dumps(out))


def parse_args(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20,
                    help="timed plain FP32 sweeps for the per-sweep statistics (minimum)")
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--target", type=float, default=1e-3)
    ap.add_argument("--max-steps", type=int, default=400)
    ap.add_argument("--subsets", type=int, default=20)
    ap.add_argument("--bp", default="ref", choices=["ref", "matched"])
    ap.add_argument("--cpu-views", type=int, default=64)
    ap.add_argument("--sweeps-only", action="store_true",
                    help="profiling: exactly --steps plain FP32 sweeps, no target, no other leg")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--form", default="both", choices=["both", "plain", "learned"])
    ap.add_argument("--weights2", default=os.path.join(ROOT, "ctrecon_amd", "weights",
                                                       "warm_start_stage2.npz"),
                    help="second-stage network (used when the file exists)")
    ap.add_argument("--after2", type=int, default=3)
    ap.add_argument("--weights", default=os.path.join(ROOT, "ctrecon_amd", "weights",
                                                      "warm_start.npz"))
    ap.add_argument("--weights-recipe", default=os.path.join(ROOT, "ctrecon_amd", "weights",
                                                             "warm_start_iter50.npz"),
                    help="network of the configured training recipe")
    ap.add_argument("--recipe", action="store_true",
                    help="also run the configured recipe's network to the target (capped at "
                         "the plain sweep count) and report it beside the shipped networks")
    ap.add_argument("--workload", default="clinical",
                    choices=["clinical", "batched", "projectors"],
                    help="clinical: the headline solve; batched: 16 uneven scans, one sweep; "
                         "projectors: every projector model on one clinical subset")
    ap.add_argument("--calib-copy", action="store_true",
                    help="only run the plain copy kernel (byte-counter calibration)")
    return ap.parse_args(argv)


def main():
    args = parse_args()
    if args.gpus != 1:
        log("bench.py measures one GPU: this build has no multi-GPU path (one GPU holds the "
            "whole volume and sinogram); run with --gpus 1.")
        sys.exit(2)

    from ctrecon_amd import geometry as G
    from ctrecon_amd import phantom as P
    from ctrecon_amd._lib import lib
    from ctrecon_amd.plan import DeviceArray, Plan, Sart, device_count

    if device_count() < 1:
        log("no GPU visible")
        sys.exit(1)

    if args.calib_copy:
        n = 1 << 26  # 64 Mi floats = 256 MiB read + 256 MiB written per launch
        a, b = DeviceArray((n,)), DeviceArray((n,))
        a.zero()
        for _ in range(5):
            lib().ctr_device_copy(b.ptr, a.ptr, n, None)
        lib().ctr_synchronize(0)
        print(json.dumps({"calib_copy_bytes_read": 4 * n, "calib_copy_bytes_written": 4 * n,
                          "launches": 5}))
        return

    if args.workload == "batched":
        batched_workload(args)
        return
    if args.workload == "projectors":
        projectors_workload(args)
        return

    cfg = G.config_clinical()
    g = G.dd3_geometry(cfg, mask=G.fov_mask(512))
    log("building phantom")
    x_true = P.ellipsoid_mixture(g, seed=1)
    plan = Plan(g, 0)
    d_true = DeviceArray.from_host(x_true)
    d_b = DeviceArray((g.PN, g.DNU, g.DNV))
    plan.fp(d_true, d_b)
    plan.synchronize()
    d_x = DeviceArray(g.vol_shape)
    m = g.mask.reshape(g.YN, g.XN).astype(bool)

    def rmse_of(d):
        x = d.get()
        return float(np.sqrt(np.mean((x - x_true)[m] ** 2)))

    have_net = os.path.exists(args.weights)
    if args.form == "learned" and not have_net:
        log(f"no network weights at {args.weights}")
        sys.exit(1)
    K = max(args.steps, 1)

    if args.sweeps_only:
        # profiling: exactly K plain FP32 sweeps from zero, nothing else
        d_x.zero()
        plan.synchronize()
        res, sweep_ms, st = plan.os_sart_run(d_x, d_b, args.subsets, K, args.bp)
        plan.synchronize()
        print(json.dumps({"sweeps": len(res), "ms_per_sweep": float(np.mean(sweep_ms)),
                          "residual_final": float(res[-1]), "setup_ms": st["setup_ms"],
                          "fp_avg_ms": st["fp_ms"] / max(st["fp_launches"], 1),
                          "bp_avg_ms": st["bp_ms"] / max(st["bp_launches"], 1)}))
        return

    run_learned = args.form in ("both", "learned") and have_net
    max_sweeps = max(args.max_steps, K)

    # ---- plain FP32: W untimed warm-up sweeps (plain launches, each one timed with HIP
    # events: the kernel averages of the roofline), then the solve from zero to the target
    # (sweeps replayed from a captured graph); its first K sweeps are the timed steps
    st = None
    if args.warmup > 0:
        d_x.zero()
        _, _, st = plan.os_sart_run(d_x, d_b, args.subsets, args.warmup, args.bp)
        plan.synchronize()
        log(f"warmup: {args.warmup} sweep(s) done")
    d_x.zero()
    plan.synchronize()
    t0 = time.perf_counter()
    res, sweep_ms, _ = plan.os_sart_run(d_x, d_b, args.subsets, max_sweeps, args.bp,
                                        args.target, stats=False)
    plan.synchronize()
    wall = time.perf_counter() - t0
    hit = np.nonzero(res <= args.target)[0]
    plain = {
        "time_to_target_s": wall if hit.size else None,
        "wall_s": wall,
        "sweeps": len(res),
        "sweeps_to_target": int(hit[0]) + 1 if hit.size else None,
        "ms_per_sweep": float(np.mean(sweep_ms)),
        "gpu_sweeps_time_to_target_s": (float(np.sum(sweep_ms[:hit[0] + 1])) / 1e3
                                        if hit.size else None),
        "residual_sweep_final": float(res[-1]),
        "residual_true_final": plan.residual(d_x, d_b),
        "rmse_vs_truth_in_mask": rmse_of(d_x),
        "timed_ms": [float(v) for v in sweep_ms[:K]],
    }
    log(f"plain FP32: {len(res)} sweeps, wall {wall:.2f} s, residual {res[-1]:.3e}, "
        f"time to target {plain['time_to_target_s']}")
    if len(res) < K:  # fewer sweeps than the timed steps: continue for the statistics only
        S = Sart(plan, d_b, args.subsets, args.bp)
        _, more_ms, _ = S.sweeps(d_x, K - len(res))
        S.close()
        plain["timed_ms"] += [float(v) for v in more_ms]
    log(f"timed steps: {len(plain['timed_ms'])} plain FP32 sweeps, "
        f"{np.mean(plain['timed_ms']):.1f} ms each")

    # ---- the learned warm start with BF16 operands, from zero to the target
    learned = learned_sino = recipe = None
    if run_learned:
        from ctrecon_amd.learned import warm_start_os_sart
        from ctrecon_amd.unet import UNet

        net = UNet.from_file(args.weights)
        stages = []
        if os.path.exists(args.weights2):
            stages.append((UNet.from_file(args.weights2), args.after2))

        def learned_leg(storage, net=net, stages=stages, weights=args.weights, cap=None):
            plan.set_storage(storage)
            # warm-up: one sweep and one network pass (code loading, network buffers)
            d_x.zero()
            plan.os_sart_run(d_x, d_b, args.subsets, 1, args.bp)
            tmp = DeviceArray(g.vol_shape)
            net.forward(d_x, d_x, tmp)
            plan.synchronize()
            del tmp
            d_x.zero()
            plan.synchronize()
            t0 = time.perf_counter()
            info = warm_start_os_sart(plan, net, d_b, d_x, args.subsets, args.target,
                                      cap or max_sweeps, args.bp, stages=stages)
            plan.synchronize()
            wall = time.perf_counter() - t0
            plan.set_storage("f32")
            r = np.array(info["residuals"])
            out = {
                "storage": storage,
                "time_to_target_s": wall if r[-1] <= args.target else None,
                "wall_s": wall,
                "sweeps": int(info["sweeps"]),
                "fallback_taken": bool(info["fallback"]),
                "later_stage_fallbacks": info["stage_fallbacks"],
                "phase_s": {k: info[k] for k in ("setup_s", "pre_sweeps_s", "network_s",
                                                 "post_sweeps_s")},
                "residual_sweep_final": float(r[-1]),
                "residual_true_final": plan.residual(d_x, d_b),
                "rmse_vs_truth_in_mask": rmse_of(d_x),
                "weights": os.path.relpath(weights, ROOT),
                "later_stages": [os.path.relpath(args.weights2, ROOT)] if stages else [],
            }
            if cap:
                out["max_sweeps"] = cap
            log(f"learned + {storage}: {info['sweeps']} sweeps, wall {wall:.2f} s, "
                f"residual {r[-1]:.3e}, fallback {info['fallback']}")
            return out

        # BF16 volume and sinogram operands (the reported form), then BF16 sinogram
        # operands with the FP32 iterate
        learned = learned_leg("bf16")
        learned_sino = learned_leg("bf16_sino")
        # the network of the configured training recipe (targets: 50-sweep OS-SART iterates,
        # 256x256x128 training grid), first stage alone, capped at the plain sweep count
        # (time_to_target_s null: not faster than plain)
        if args.recipe and os.path.exists(args.weights_recipe):
            recipe = learned_leg("bf16", net=UNet.from_file(args.weights_recipe), stages=[],
                                 weights=args.weights_recipe,
                                 cap=plain["sweeps_to_target"] or max_sweeps)

    # ---- roofline of the dominant kernel (the back-projection update)
    if st is None:  # no warm-up sweeps: time two sweeps' launches now
        d_x.zero()
        _, _, st = plan.os_sart_run(d_x, d_b, args.subsets, 2, args.bp)
    plain["setup_ms"] = st["setup_ms"]
    fp_avg = st["fp_ms"] / max(st["fp_launches"], 1)
    bp_avg = st["bp_ms"] / max(st["bp_launches"], 1)
    cone = plan.cone_voxels()
    frame = g.DNU * g.DNV
    bytes_per_view = 4.0 * cone + 4.0 * frame
    per_launch_bytes = float(bytes_per_view.sum()) / args.subsets
    views_per_launch = g.PN / args.subsets
    dominant = "bp_win_kernel" if args.bp == "ref" else "bp_matched_win_kernel"
    achieved = per_launch_bytes / (bp_avg / 1e3)
    counters = {}
    tpath = os.path.join(ROOT, "profiles", "traffic.json")
    if os.path.exists(tpath):
        counters = json.load(open(tpath)).get(dominant + "_counters", {})
    roof = {
        "kernel": dominant,
        # the named roofline of the projectors (no MFMA in them); the counters below say
        # which unit limits the kernel
        "bound": "hbm",
        "achieved": achieved / 1e9,
        "peak": HBM_PEAK / 1e9,
        "unit": "GB/s",
        "frac": achieved / HBM_PEAK,
        "traffic": counters.get("hbm_bytes_per_launch"),
        "traffic_source": counters.get("source"),
        "limiter": counters.get("limiter"),
        "valu_issue_frac": counters.get("valu_issue_frac"),
        "salu_issue_frac": counters.get("salu_issue_frac"),
        "algorithmic_bytes_per_launch": per_launch_bytes,
        # compulsory I/O of one update launch: the subset's residual sinogram in, the
        # iterate and its column sums in, the iterate out
        "compulsory_bytes_per_launch": (4.0 * frame * views_per_launch
                                        + 3 * 4.0 * g.XN * g.YN * g.ZN),
        "avg_launch_ms": bp_avg,
        "launches_timed": st["bp_launches"],
    }
    kernels = {"fp_fused": fp_avg, "bp_update": bp_avg}
    pps = {
        "os_sart_loop": g.PN / (plain["ms_per_sweep"] / 1e3),
        "fp_kernel": views_per_launch / (fp_avg / 1e3),
        "bp_kernel": views_per_launch / (bp_avg / 1e3),
    }
    cpu = None
    if not args.no_cpu:
        log("cpu baseline sample")
        cpu = cpu_baseline(g, x_true, args.cpu_views, plain["sweeps_to_target"])
    out = headline_record(args, plain, learned, learned_sino, roof, pps, kernels, cpu,
                          have_net, g.PN, recipe)
    print(json.dumps(out))


if __name__ == "__main__":
    main()

"""Learned warm start for OS-SART: training data, training, and the warm-started solve.

Pipeline (all projector and network work on the GPU through the C ABI):
  1. x1, x2 = the OS-SART iterates after one and two sweeps from x0 = 0, and
     delta = x2 - x1 (the back-projected, normalised residual of sweep 2);
  2. x3 = x2 + net(x2, delta), the network's predicted update;
  3. plain OS-SART from x3 until the data residual reaches the target.  If the first
     sweep from x3 sees a residual that is not below the one seen during sweep 2, the
     iterate is reset to x2 and plain OS-SART continues from there (the fallback);
  4. optionally later networks at the start of each slow tail: stage k acts m_k sweeps
     after the previous network's update, x <- x_m + net_k(x_m, x_m - x_(m-1)), with the
     same fallback.
Training pairs come from plain OS-SART itself: on seeded phantoms the input is
(x2, delta) and the target update is x_K - x2 (x_K the K-sweep iterate) or the phantom
minus x2; the pairs of a later stage come from the warm-started solve (with the stages
before it) at the point where it acts.

    python -m ctrecon_amd.learned train --out weights.npz [--phantoms 64 --iters 50 ...]
    python -m ctrecon_amd.learned train --out stage2.npz --stage-from weights.npz \
        --stage-after 3 --targets truth ...
    python -m ctrecon_amd.learned train --out stage3.npz \
        --stage-from weights.npz,stage2.npz --stage-afters 3 --stage-after 2 ...
"""

from __future__ import annotations

import argparse
import math
import sys
import time

import numpy as np

from . import geometry as G
from . import phantom as PH
from .plan import DeviceArray, Plan, Sart
from ._lib import check, lib
from .unet import UNet, init_params

__all__ = ["perlin3", "training_phantom", "make_pair", "make_pair_stage", "train",
           "warm_start_os_sart"]


def _log(*a):
    print(*a, file=sys.stderr, flush=True)


def perlin3(shape, period, rng, step=1) -> np.ndarray:
    """Gradient (Perlin) noise on a grid: random unit gradients on a lattice of spacing
    ``period`` voxels, quintic fade, trilinear blend of the corner dot products.
    Values lie in about [-1, 1].  step > 1 evaluates it every ``step`` voxels and
    repeats the values (large grids)."""
    if step > 1:
        coarse = perlin3(tuple(-(-n // step) for n in shape), period / step, rng)
        full = coarse.repeat(step, 0).repeat(step, 1).repeat(step, 2)
        return np.ascontiguousarray(full[:shape[0], :shape[1], :shape[2]])
    Y, X, Z = shape
    n = [int(math.ceil(s / period)) + 2 for s in shape]
    gr = rng.normal(size=(n[0], n[1], n[2], 3)).astype(np.float32)
    gr /= np.linalg.norm(gr, axis=-1, keepdims=True)

    def axis(m):
        f = np.arange(m, dtype=np.float32) / period
        i = np.floor(f).astype(np.int64)
        return i, (f - i).astype(np.float32)

    (iy, ty), (ix, tx), (iz, tz) = axis(Y), axis(X), axis(Z)
    fade = lambda t: t * t * t * (t * (t * 6 - 15) + 10)  # noqa: E731
    fy, fx, fz = fade(ty), fade(tx), fade(tz)
    out = np.zeros(shape, np.float32)
    for cy in (0, 1):
        wy = (fy if cy else 1 - fy)[:, None, None]
        for cx in (0, 1):
            wx = (fx if cx else 1 - fx)[None, :, None]
            for cz in (0, 1):
                wz = (fz if cz else 1 - fz)[None, None, :]
                Gc = gr[(iy + cy)[:, None, None], (ix + cx)[None, :, None], (iz + cz)[None, None, :]]
                dot = (Gc[..., 0] * (ty - cy)[:, None, None] + Gc[..., 1] * (tx - cx)[None, :, None]
                       + Gc[..., 2] * (tz - cz)[None, None, :])
                out += wy * wx * wz * dot
    return out * math.sqrt(2.0)


def training_phantom(g, seed: int) -> np.ndarray:
    """Ellipsoid mixture (densities 0-2; 40 ellipsoids of 5-60 mm in-plane / 5-40 mm in z
    on clinical-size grids, 14 of 4-45 / 4-30 mm on smaller ones) plus a two-octave
    Perlin texture inside the support, clipped to [0, 2]; float32 vol[y, x, z]."""
    big = g.XN * g.YN * g.ZN > (1 << 25)
    if big:
        vol = PH.ellipsoid_mixture(g, seed=seed, n=40, texture_sigma=0.01)
    else:
        vol = PH.ellipsoid_mixture(g, seed=seed, n=14, axes_xy=(4.0, 45.0),
                                   axes_z=(4.0, 30.0), texture_sigma=0.01)
    rng = np.random.default_rng(seed + 7919)
    st = 2 if big else 1
    tex = 0.08 * perlin3(g.vol_shape, 24, rng, st) + 0.04 * perlin3(g.vol_shape, 8, rng, st)
    vol = np.where(vol > 0, vol + tex * np.maximum(vol, 0.25), 0.0)
    return np.clip(vol, 0.0, 2.0).astype(np.float32)


def _sub(d_out, d_a, d_b):
    check(lib().ctr_device_sub(d_out.ptr, d_a.ptr, d_b.ptr, d_out.nbytes // 4, None))


def _axpy(d_y, d_x, a):
    check(lib().ctr_device_axpy(d_y.ptr, d_x.ptr, float(a), d_y.nbytes // 4, None))


def make_pair(plan: Plan, vol, iters=50, osNum=20, kind="ref"):
    """(x2, delta, x_K - x2, vol - x2) float32 host arrays from plain OS-SART on
    b = A vol: the network input and two targets (the K-sweep iterate, the phantom)."""
    g = plan.g
    d_b = DeviceArray(g.prj_shape)
    plan.fp(DeviceArray.from_host(vol), d_b)
    S = Sart(plan, d_b, osNum, kind)
    dx = DeviceArray(g.vol_shape)
    dx.zero()
    S.sweeps(dx, 1)
    x1 = dx.get()
    S.sweeps(dx, 1)
    x2 = dx.get()
    if iters > 2:  # the K-sweep target (not needed when training on the phantom only)
        S.sweeps(dx, iters - 2)
    xk = dx.get()
    S.close()
    return (x2, x2 - x1, xk - x2 if iters > 2 else None,
            np.asarray(vol, np.float32) - x2)


def _net_update(plan: Plan, net: UNet, d_x: DeviceArray, d_prev: DeviceArray,
                d_keep: DeviceArray):
    """d_keep <- d_x; d_x <- d_x + net(d_x, d_x - d_prev) (the update masked to the FOV)."""
    shape = d_x.shape
    delta = DeviceArray(shape)
    _sub(delta, d_x, d_prev)
    upd = DeviceArray(shape)
    net.forward(d_x, delta, upd)
    plan.mask_volume(upd)
    d_keep.copy_from(d_x)
    _axpy(d_x, upd, 1.0)


def make_pair_stage(plan: Plan, vol, net1: UNet, stages=(), after=3, osNum=20, kind="ref"):
    """(x_m, x_m - x_(m-1), None, vol - x_m) float32 host arrays: the warm-started solve on
    b = A vol (two sweeps, net1, then for each (net_k, m_k) of `stages` m_k sweeps and
    net_k's update, then `after` sweeps) at the point where the next network acts."""
    g = plan.g
    d_b = DeviceArray(g.prj_shape)
    plan.fp(DeviceArray.from_host(vol), d_b)
    S = Sart(plan, d_b, osNum, kind)
    dx = DeviceArray(g.vol_shape)
    dx.zero()
    prev = DeviceArray(g.vol_shape)
    keep = DeviceArray(g.vol_shape)
    S.sweeps(dx, 1)
    prev.copy_from(dx)
    S.sweeps(dx, 1)
    _net_update(plan, net1, dx, prev, keep)
    for netk, mk in list(stages) + [(None, after)]:
        if mk > 1:
            S.sweeps(dx, mk - 1)
        prev.copy_from(dx)
        S.sweeps(dx, 1)
        if netk is not None:
            _net_update(plan, netk, dx, prev, keep)
    x = dx.get()
    dprev = prev.get()
    S.close()
    return (x, x - dprev, None, np.asarray(vol, np.float32) - x)


def _cosine(step, steps, lr):
    return lr * 0.5 * (1.0 + math.cos(math.pi * min(step, steps) / steps))


def running_std(arrays) -> float:
    """Standard deviation over all elements of several arrays, accumulated array by array
    in FP64 (no concatenated copy of a clinical-size training set)."""
    n, s1, s2 = 0, 0.0, 0.0
    for a in arrays:
        a64 = a.astype(np.float64, copy=False).ravel()
        n += a64.size
        s1 += float(a64.sum())
        s2 += float(np.dot(a64, a64))
    mean = s1 / n
    return float(np.sqrt(max(s2 / n - mean * mean, 0.0)))


def train(pairs, steps=2000, crop=(128, 128, 64), lr=1e-3, seed=0, val=None, log=_log):
    """Adam on random crops of the training pairs; returns the trained UNet."""
    s_d = 1.0 / max(running_std(p[1] for p in pairs), 1e-12)
    s_out = max(running_std(p[2] for p in pairs), 1e-12)
    net = UNet(init_params(seed), (1.0, s_d, s_out))
    rng = np.random.default_rng(seed)
    dx, dd, dt = (DeviceArray(crop) for _ in range(3))
    t0 = time.time()
    hist = []
    for step in range(steps):
        x2, de, tg = pairs[rng.integers(len(pairs))]
        o = [int(rng.integers(0, s - c + 1)) for s, c in zip(x2.shape, crop)]
        sl = tuple(slice(a, a + c) for a, c in zip(o, crop))
        dx.upload(np.ascontiguousarray(x2[sl]))
        dd.upload(np.ascontiguousarray(de[sl]))
        dt.upload(np.ascontiguousarray(tg[sl]))
        hist.append(net.train_step(dx, dd, dt, _cosine(step, steps, lr)))
        if (step + 1) % max(1, steps // 20) == 0:
            log(f"  step {step + 1}/{steps} loss {np.mean(hist[-50:]):.4e} "
                f"({time.time() - t0:.0f} s)")
    if val:
        log(f"  validation: relative error of the predicted update {evaluate(net, val):.4f}")
    return net


def evaluate(net: UNet, pairs) -> float:
    """Mean over pairs of ||net(x2, delta) - target|| / ||target||."""
    errs = []
    for x2, de, tg in pairs:
        out = DeviceArray(x2.shape)
        net.forward(DeviceArray.from_host(x2), DeviceArray.from_host(de), out)
        u = out.get()
        errs.append(float(np.linalg.norm(u - tg) / np.linalg.norm(tg)))
    return float(np.mean(errs))


def warm_start_os_sart(plan: Plan, net: UNet, d_b: DeviceArray, d_x: DeviceArray, osNum=20,
                       stop_resid=1e-3, max_sweeps=400, kind="ref", fallback=True, sart=None,
                       stages=()):
    """The learned warm start followed by plain OS-SART (module docstring); `stages` lists
    the later networks as (net_k, m_k), net_k acting m_k sweeps after the previous
    network's update.  d_x receives the result.  Returns a dict with the per-sweep
    residuals, the number of sweeps, whether the first network's fallback was taken and,
    per later stage, whether its fallback was taken (or None: not reached), and the wall
    time of each phase (s)."""
    t0 = time.perf_counter()
    S = sart if sart is not None else Sart(plan, d_b, osNum, kind)
    t_setup = time.perf_counter() - t0
    g = plan.g
    shape = g.vol_shape
    d_x.zero()
    res = list(S.sweeps(d_x, 1)[0])
    x1 = DeviceArray(shape)
    x1.copy_from(d_x)
    res += list(S.sweeps(d_x, 1)[0])
    t1 = time.perf_counter()
    delta = DeviceArray(shape)
    _sub(delta, d_x, x1)
    net.forward(d_x, delta, x1)  # x1 now holds the predicted update
    plan.mask_volume(x1)  # no update outside the FOV mask (those voxels stay 0)
    delta.copy_from(d_x)  # delta now holds x2 (kept for the fallback)
    _axpy(d_x, x1, 1.0)
    plan.synchronize()
    t2 = time.perf_counter()
    took_fallback = False
    r3 = S.sweeps(d_x, 1, stop_resid)[0]
    if fallback and r3[0] >= res[-1]:
        d_x.copy_from(delta)
        took_fallback = True
        res_after = list(r3)  # the wasted sweep is counted
    else:
        res_after = list(r3)
    done = 3
    reached = r3[-1] <= stop_resid and not took_fallback
    stage_fallbacks = [None] * len(stages)
    net2_s = 0.0
    prev = DeviceArray(shape)
    keep = DeviceArray(shape)
    for k, (netk, mk) in enumerate(stages):
        if took_fallback or reached or any(stage_fallbacks[:k]):
            break
        # m_k sweeps from the previous update in all (one done), the last one's update kept
        if mk > 2:
            r = S.sweeps(d_x, mk - 2, stop_resid)[0]
            res_after += list(r)
            done += len(r)
            reached = r[-1] <= stop_resid
            if reached:
                break
        prev.copy_from(d_x)
        r = S.sweeps(d_x, 1, stop_resid)[0]
        res_after += list(r)
        done += 1
        reached = r[-1] <= stop_resid
        if reached:
            break
        tn = time.perf_counter()
        _net_update(plan, netk, d_x, prev, keep)
        plan.synchronize()
        net2_s += time.perf_counter() - tn
        r5 = S.sweeps(d_x, 1, stop_resid)[0]
        done += 1
        res_after += list(r5)
        stage_fallbacks[k] = bool(fallback and r5[0] >= res_after[-2])
        if stage_fallbacks[k]:
            d_x.copy_from(keep)
        reached = r5[-1] <= stop_resid and not stage_fallbacks[k]
    if not reached and max_sweeps > done:
        r = S.sweeps(d_x, max_sweeps - done, stop_resid)[0]
        res_after += list(r)
        done += len(r)
    t3 = time.perf_counter()
    if sart is None:
        S.close()
    return {"residuals": res + res_after, "sweeps": done, "fallback": took_fallback,
            "stage_fallbacks": stage_fallbacks, "setup_s": t_setup,
            "pre_sweeps_s": t1 - t0 - t_setup, "network_s": t2 - t1 + net2_s,
            "post_sweeps_s": t3 - t2 - net2_s, "total_s": t3 - t0}


def _phantom_job(nxy, nz, seed):
    g = G.dd3_geometry(G.config_training(nxy, nz), mask=G.fov_mask(nxy))
    return training_phantom(g, seed)


def compare(plan: Plan, net: UNet, vol, target=1e-3, osNum=20, max_sweeps=400, kind="ref",
            plain_storage=None, warm_storage=None, stages=()):
    """Plain OS-SART vs the warm start on b = A vol (storage modes default to the
    plan's): sweeps and wall times to the residual target."""
    g = plan.g
    keep = getattr(plan, "storage", "f32")
    d_b = DeviceArray(g.prj_shape)
    plan.set_storage("f32")
    plan.fp(DeviceArray.from_host(vol), d_b)
    dx = DeviceArray(g.vol_shape)
    dx.zero()
    plan.set_storage(plain_storage or keep)
    plan.synchronize()
    t0 = time.perf_counter()
    S = Sart(plan, d_b, osNum, kind)
    r_plain = S.sweeps(dx, max_sweeps, target)[0]
    plan.synchronize()
    t_plain = time.perf_counter() - t0
    S.close()
    plan.set_storage(warm_storage or keep)
    info = warm_start_os_sart(plan, net, d_b, dx, osNum, target, max_sweeps, kind,
                              stages=stages)
    plan.set_storage(keep)
    return {"plain_sweeps": len(r_plain), "plain_s": t_plain, "plain_final": float(r_plain[-1]),
            "warm_sweeps": info["sweeps"], "warm_s": info["total_s"],
            "warm_final": float(info["residuals"][-1]), "fallback": info["fallback"],
            "stage_fallbacks": info["stage_fallbacks"],
            "residual_after_warm_start": float(info["residuals"][2]),
            "residual_plain_sweep3": float(r_plain[2]) if len(r_plain) > 2 else None}


def _eval(a):
    g = G.dd3_geometry(G.config_training(a.nxy, a.nz), mask=G.fov_mask(a.nxy))
    plan = Plan(g, 0)
    net = UNet.from_file(a.weights)
    stages = _stage_list(a.stages, a.afters)
    for seed in a.seeds:
        vol = PH.ellipsoid_mixture(g, seed=seed) if a.mixture else training_phantom(g, seed)
        r = compare(plan, net, vol, a.target, a.subsets, a.max_sweeps,
                    plain_storage=a.plain_storage, warm_storage=a.warm_storage, stages=stages)
        print(f"seed {seed}: " + ", ".join(f"{k} {v:.4g}" if isinstance(v, float) else f"{k} {v}"
                                           for k, v in r.items()), flush=True)


def _stage_list(files, afters):
    """[(UNet, m_k)] from a comma list of weights files and one of sweep counts."""
    fs = [f for f in (files or "").split(",") if f]
    ms = [int(m) for m in (afters or "").split(",") if m]
    if len(fs) != len(ms):
        raise SystemExit("one sweep count per later-stage weights file")
    return [(UNet.from_file(f), m) for f, m in zip(fs, ms)]


def main(argv=None):
    ap = argparse.ArgumentParser(prog="python -m ctrecon_amd.learned")
    sub = ap.add_subparsers(dest="cmd", required=True)
    tr = sub.add_parser("train", help="generate pairs with plain OS-SART and train")
    tr.add_argument("--out", required=True, help="weights file; with several targets, "
                    "one file per target: <out>_<target>.npz")
    tr.add_argument("--targets", default="iter", help="comma list of 'iter' (the K-sweep "
                    "OS-SART iterate) and 'truth' (the phantom)")
    tr.add_argument("--eval-seeds", type=int, nargs="*", default=[],
                    help="held-out training-recipe phantoms to evaluate on")
    tr.add_argument("--eval-mixture-seeds", type=int, nargs="*", default=[],
                    help="held-out plain ellipsoid mixtures (the benchmark's recipe)")
    tr.add_argument("--target", type=float, default=1e-3)
    tr.add_argument("--phantoms", type=int, default=64)
    tr.add_argument("--val", type=int, default=4)
    tr.add_argument("--seed0", type=int, default=1000)
    tr.add_argument("--iters", type=int, default=50)
    tr.add_argument("--subsets", type=int, default=20)
    tr.add_argument("--steps", type=int, default=2000)
    tr.add_argument("--lr", type=float, default=1e-3)
    tr.add_argument("--nxy", type=int, default=256)
    tr.add_argument("--nz", type=int, default=128)
    tr.add_argument("--storage", default="bf16")
    tr.add_argument("--workers", type=int, default=12)
    tr.add_argument("--stage-from", default=None,
                    help="comma list: the first network's weights, then those of the later "
                         "stages already trained; trains the next stage on the "
                         "warm-started solve")
    tr.add_argument("--stage-afters", default="",
                    help="comma list: sweeps before each later stage already trained")
    tr.add_argument("--stage-after", type=int, default=3,
                    help="sweeps before the stage being trained")
    ev = sub.add_parser("eval", help="sweeps to a residual target, plain vs warm-started")
    ev.add_argument("--weights", required=True)
    ev.add_argument("--seeds", type=int, nargs="+", default=[20000, 20001])
    ev.add_argument("--target", type=float, default=1e-3)
    ev.add_argument("--subsets", type=int, default=20)
    ev.add_argument("--nxy", type=int, default=256)
    ev.add_argument("--nz", type=int, default=128)
    ev.add_argument("--max-sweeps", type=int, default=400)
    ev.add_argument("--mixture", action="store_true",
                    help="plain 40-ellipsoid mixtures (the benchmark's phantom recipe)")
    ev.add_argument("--plain-storage", default="f32")
    ev.add_argument("--warm-storage", default="bf16")
    ev.add_argument("--stages", default="", help="comma list of later-stage weights")
    ev.add_argument("--afters", default="", help="comma list: sweeps before each")
    a = ap.parse_args(argv)
    if a.cmd == "eval":
        return _eval(a)
    t0 = time.time()
    # phantoms on the host cores, in fresh processes started before this one touches
    # the GPU
    seeds = [a.seed0 + i for i in range(a.phantoms + a.val)]
    import multiprocessing as mp

    with mp.get_context("spawn").Pool(max(1, min(a.workers, len(seeds)))) as pool:
        vols = pool.starmap(_phantom_job, [(a.nxy, a.nz, s) for s in seeds])
    _log(f"{len(vols)} phantoms ({time.time() - t0:.0f} s)")
    cfg = G.config_training(a.nxy, a.nz)
    g = G.dd3_geometry(cfg, mask=G.fov_mask(a.nxy))
    plan = Plan(g, 0)
    plan.set_storage(a.storage)
    pairs = []
    net1, before = None, []
    if a.stage_from:
        files = a.stage_from.split(",")
        net1 = UNet.from_file(files[0])
        before = _stage_list(",".join(files[1:]), a.stage_afters)
    for i, vol in enumerate(vols):
        if net1 is not None:
            pairs.append(make_pair_stage(plan, vol, net1, before, a.stage_after, a.subsets))
        else:
            pairs.append(make_pair(plan, vol, a.iters, a.subsets))
        _log(f"pair {i + 1}/{len(vols)} ({time.time() - t0:.0f} s)")
    del vols
    targets = [t.strip() for t in a.targets.split(",") if t.strip()]
    nets = {}
    for tname in targets:
        ti = {"iter": 2, "truth": 3}[tname]
        sel = [(p[0], p[1], p[ti]) for p in pairs]
        _log(f"training on target '{tname}'")
        net = train(sel[:a.phantoms], a.steps, lr=a.lr, val=sel[a.phantoms:])
        out = a.out if len(targets) == 1 else a.out.replace(".npz", f"_{tname}.npz")
        net.save(out)
        nets[tname] = net
        _log(f"saved {out} ({time.time() - t0:.0f} s)")
    del pairs
    evals = [(s_, lambda s_=s_: training_phantom(g, s_)) for s_ in a.eval_seeds]
    evals += [(s_, lambda s_=s_: PH.ellipsoid_mixture(g, seed=s_)) for s_ in a.eval_mixture_seeds]
    for seed, make in evals:
        vol = make()
        for tname, net in nets.items():
            if net1 is not None:
                r = compare(plan, net1, vol, a.target, a.subsets,
                            stages=before + [(net, a.stage_after)])
            else:
                r = compare(plan, net, vol, a.target, a.subsets)
            print(f"eval seed {seed} target {tname}: " + ", ".join(
                f"{k} {v:.4g}" if isinstance(v, float) else f"{k} {v}" for k, v in r.items()),
                flush=True)


if __name__ == "__main__":
    main()

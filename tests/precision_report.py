"""Precision report of the BF16-storage form against plain FP32, both measured against
the FP64 oracle, on the small helical case (64^3 Shepp-Logan at 1 mm, 256x16 detector,
720 views): relative L2 error per projection, per back projection, per OS-SART
(10 subsets) and CG iterate, and of the final volumes.

    python tests/precision_report.py [--iters 30] [--seed 0]   (needs a GPU; prints markdown)
"""

from __future__ import annotations

import argparse
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
sys.path.insert(0, os.path.dirname(HERE))

from bf16 import quantize  # noqa: E402
from cases import small  # noqa: E402


def rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def per_view(a, b):
    return np.array([rel(a[k], b[k]) for k in range(a.shape[0]) if np.any(b[k])])


def main():
    from ctrecon_amd import phantom as P
    from ctrecon_amd.plan import DeviceArray, Plan
    from oracle import oracle as O

    ap = argparse.ArgumentParser()
    ap.add_argument("--iters", type=int, default=30)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--subsets", type=int, default=10)
    a = ap.parse_args()
    cfg, g = small()
    vol = P.shepp_logan(g, seed=a.seed).astype(np.float32)
    plan = Plan(g, 0)
    t0 = time.time()
    lines = []
    say = lines.append

    def log(msg):
        print(msg, file=sys.stderr, flush=True)

    # ---- projections ----
    ref_fp = O.fp(g, vol, "direct")
    f32 = plan.forward(vol)
    b16 = plan.forward_bf16(vol)
    y = ref_fp.astype(np.float32)
    ref_bp = O.bp(g, y, "direct")
    bp32 = plan.back(y)
    bp16 = plan.back_bf16(y)
    log(f"projections done ({time.time() - t0:.0f} s)")
    say("| quantity | plain FP32 | BF16 storage |")
    say("|---|---|---|")
    say(f"| forward projection, all views | {rel(f32, ref_fp):.2e} | {rel(b16, ref_fp):.2e} |")
    pv32, pv16 = per_view(f32, ref_fp), per_view(b16, ref_fp)
    say(f"| forward projection, worst view | {pv32.max():.2e} | {pv16.max():.2e} |")
    say(f"| back projection (reference), all views | {rel(bp32, ref_bp):.2e} | "
        f"{rel(bp16, ref_bp):.2e} |")
    say(f"| BF16 rounding of the volume itself | - | {rel(quantize(vol), vol):.2e} |")

    # ---- OS-SART, one sweep per call (the iterate is the whole state) ----
    b = ref_fp.astype(np.float32)
    db = DeviceArray.from_host(b)
    xr, res_r, tr = O.os_sart(g, b, osNum=a.subsets, iterNum=a.iters, trace=True)
    log(f"oracle OS-SART done ({time.time() - t0:.0f} s)")
    sart = {}
    for storage in ("f32", "bf16"):
        plan.set_storage(storage)
        dx = DeviceArray(g.vol_shape)
        dx.zero()
        errs, res = [], []
        for it in range(a.iters):
            res.append(float(plan.os_sart(dx, db, a.subsets, 1)[0]))
            errs.append(rel(dx.get(), tr[it]))
        sart[storage] = (np.array(errs), np.array(res), dx.get())
    plan.set_storage("f32")
    # ---- CG: iterate k from a fresh start with k iterations ----
    xc, resc, trc = O.conjgrad(g, b, iterNum=a.iters, trace=True)
    log(f"oracle CG done ({time.time() - t0:.0f} s)")
    cg = {}
    for storage in ("f32", "bf16"):
        plan.set_storage(storage)
        errs = []
        last = None
        for k in range(1, a.iters + 1):
            dx = DeviceArray(g.vol_shape)
            dx.zero()
            plan.conjgrad(dx, db, k)
            last = dx.get()
            errs.append(rel(last, trc[k - 1]))
        cg[storage] = (np.array(errs), last)
    plan.set_storage("f32")
    for name, d in (("OS-SART", sart), ("CG", cg)):
        e32, e16 = d["f32"][0], d["bf16"][0]
        say(f"| {name} iterate, max over {a.iters} iterations | {e32.max():.2e} | {e16.max():.2e} |")
        say(f"| {name} final volume (iteration {a.iters}) | {e32[-1]:.2e} | {e16[-1]:.2e} |")
    m = g.mask.reshape(g.YN, g.XN)[:, :, None].astype(bool) & np.ones(g.vol_shape, bool)
    for name, x64, d in (("OS-SART", xr, sart), ("CG", xc, cg)):
        r64 = float(np.sqrt(np.mean((x64 - vol)[m] ** 2)))
        r32 = float(np.sqrt(np.mean((d["f32"][-1] - vol)[m] ** 2)))
        r16 = float(np.sqrt(np.mean((d["bf16"][-1] - vol)[m] ** 2)))
        say(f"| {name} RMSE vs phantom in mask (FP64 oracle {r64:.4e}) | {r32:.4e} | {r16:.4e} |")
    say(f"| OS-SART data residual after {a.iters} sweeps (FP64 oracle {res_r[-1]:.3e}) | "
        f"{sart['f32'][1][-1]:.3e} | {sart['bf16'][1][-1]:.3e} |")
    say("")
    say("Per-iterate relative L2 error vs the FP64 oracle iterate:")
    say("")
    say("| iteration | OS-SART FP32 | OS-SART BF16 | CG FP32 | CG BF16 |")
    say("|---|---|---|---|---|")
    for it in sorted({0, 1, 2, 4, 9, 14, 19, 24, a.iters - 1}):
        if it < a.iters:
            say(f"| {it + 1} | {sart['f32'][0][it]:.2e} | {sart['bf16'][0][it]:.2e} | "
                f"{cg['f32'][0][it]:.2e} | {cg['bf16'][0][it]:.2e} |")
    print("\n".join(lines), flush=True)
    log(f"total {time.time() - t0:.0f} s")


if __name__ == "__main__":
    main()

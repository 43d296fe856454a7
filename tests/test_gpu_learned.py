"""What the trained warm-start network gains over plain OS-SART.

A held-out phantom (seed outside the training seeds) on the training geometry (clinical
detector and trajectory, 256x256x128 at 0.7 mm), BF16 storage: plain OS-SART and the
warm-started solve (two sweeps, the network's update, plain sweeps, fallback) run to the
same data-residual target; the warm start must need fewer sweeps and must not take the
fallback, and its result must stay as close to the phantom as the plain one."""

import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

WEIGHTS = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                       "ctrecon_amd", "weights", "warm_start.npz")


def test_warm_start_saves_sweeps(gpu_plan_factory):
    from ctrecon_amd import geometry as G
    from ctrecon_amd import learned as L
    from ctrecon_amd.plan import DeviceArray, Plan, Sart
    from ctrecon_amd.unet import UNet

    print("[learned warm start] seed 30000, 256x256x128, target 2e-3", flush=True)
    assert os.path.exists(WEIGHTS), "trained weights are part of the product"
    g = G.dd3_geometry(G.config_training(), mask=G.fov_mask(256))
    vol = L.training_phantom(g, 30000)
    plan = Plan(g, 0)
    plan.set_storage("bf16")
    net = UNet.from_file(WEIGHTS)
    d_b = DeviceArray(g.prj_shape)
    plan.fp(DeviceArray.from_host(vol), d_b)
    target = 2e-3
    dx = DeviceArray(g.vol_shape)
    dx.zero()
    S = Sart(plan, d_b, 20)
    r_plain = S.sweeps(dx, 300, target)[0]
    x_plain = dx.get()
    info = L.warm_start_os_sart(plan, net, d_b, dx, 20, target, 300, sart=S)
    x_warm = dx.get()
    m = g.mask.reshape(g.YN, g.XN).astype(bool)
    rmse_p = float(np.sqrt(np.mean((x_plain - vol)[m] ** 2)))
    rmse_w = float(np.sqrt(np.mean((x_warm - vol)[m] ** 2)))
    print(f"  plain: {len(r_plain)} sweeps (final {r_plain[-1]:.3e}, RMSE {rmse_p:.4f}); "
          f"warm: {info['sweeps']} sweeps (final {info['residuals'][-1]:.3e}, RMSE {rmse_w:.4f}, "
          f"fallback {info['fallback']}); residual after the warm start "
          f"{info['residuals'][2]:.3e} vs plain sweep 3 {r_plain[2]:.3e}", flush=True)
    assert r_plain[-1] <= target and info["residuals"][-1] <= target
    assert not info["fallback"]
    assert info["sweeps"] < len(r_plain)
    assert rmse_w <= 1.01 * rmse_p


WEIGHTS2 = os.path.join(os.path.dirname(WEIGHTS), "warm_start_stage2.npz")


def test_second_stage_saves_sweeps(gpu_plan_factory):
    """The second-stage network (applied after 3 sweeps from the warm start) against the
    first stage alone, at clinical size on a held-out phantom of the benchmark's recipe
    (40-ellipsoid mixture, seed 10002), BF16 storage, to rel. residual 1e-3: fewer sweeps,
    no fallback, RMSE to the phantom within 1% of the first stage's or better."""
    from ctrecon_amd import geometry as G
    from ctrecon_amd import learned as L
    from ctrecon_amd import phantom as PH
    from ctrecon_amd.plan import DeviceArray, Plan, Sart
    from ctrecon_amd.unet import UNet

    print("[second stage] seed 10002, 512x512x256, target 1e-3", flush=True)
    assert os.path.exists(WEIGHTS2), "trained weights are part of the product"
    g = G.dd3_geometry(G.config_clinical(), mask=G.fov_mask(512))
    vol = PH.ellipsoid_mixture(g, seed=10002)
    plan = Plan(g, 0)
    d_b = DeviceArray(g.prj_shape)
    plan.fp(DeviceArray.from_host(vol), d_b)
    plan.set_storage("bf16")
    net, net2 = UNet.from_file(WEIGHTS), UNet.from_file(WEIGHTS2)
    target = 1e-3
    m = g.mask.reshape(g.YN, g.XN).astype(bool)
    S = Sart(plan, d_b, 20)
    dx = DeviceArray(g.vol_shape)
    one = L.warm_start_os_sart(plan, net, d_b, dx, 20, target, 200, sart=S)
    rmse1 = float(np.sqrt(np.mean((dx.get() - vol)[m] ** 2)))
    two = L.warm_start_os_sart(plan, net, d_b, dx, 20, target, 200, sart=S, stages=[(net2, 3)])
    rmse2 = float(np.sqrt(np.mean((dx.get() - vol)[m] ** 2)))
    print(f"  first stage: {one['sweeps']} sweeps (final {one['residuals'][-1]:.3e}, RMSE "
          f"{rmse1:.4f}); both stages: {two['sweeps']} sweeps (final {two['residuals'][-1]:.3e}, "
          f"RMSE {rmse2:.4f}, fallback {two['fallback']}/{two['stage_fallbacks']})", flush=True)
    assert one["residuals"][-1] <= target and two["residuals"][-1] <= target
    assert not two["fallback"] and two["stage_fallbacks"] == [False]
    assert two["sweeps"] < one["sweeps"]
    assert rmse2 <= 1.01 * rmse1


def _harmful(net_path):
    """The trained network with its output negated and scaled x10: an update that raises
    the data residual, which must trigger the fallback."""
    from ctrecon_amd.unet import UNet

    net = UNet.from_file(net_path)
    s_x, s_d, s_out = net.scales
    net.set_scales(s_x, s_d, -10.0 * s_out)
    return net


def test_first_stage_fallback_continues_plain_from_x2(gpu_plan_factory):
    """A harmful first-stage update: the fallback is taken, the iterate is reset to x2 and
    the solve is plain OS-SART from x2 (one wasted sweep), so its residuals and its result
    are those of a plain session continued from x2."""
    from ctrecon_amd import geometry as G
    from ctrecon_amd import learned as L
    from ctrecon_amd.plan import DeviceArray, Plan, Sart

    print("[fallback, first stage] seed 30001, 256x256x128, bf16, target 3e-3", flush=True)
    g = G.dd3_geometry(G.config_training(), mask=G.fov_mask(256))
    vol = L.training_phantom(g, 30001)
    plan = Plan(g, 0)
    plan.set_storage("bf16")
    d_b = DeviceArray(g.prj_shape)
    plan.fp(DeviceArray.from_host(vol), d_b)
    target = 3e-3
    S = Sart(plan, d_b, 20)
    dx = DeviceArray(g.vol_shape)
    info = L.warm_start_os_sart(plan, _harmful(WEIGHTS), d_b, dx, 20, target, 300, sart=S)
    x_fb = dx.get()
    dx.zero()
    r_plain = S.sweeps(dx, 300, target)[0]
    x_plain = dx.get()
    r = np.asarray(info["residuals"])
    print(f"  fallback {info['fallback']}: {info['sweeps']} sweeps, residual after the harmful "
          f"update {r[2]:.3e} (sweep 2: {r[1]:.3e}); plain {len(r_plain)} sweeps", flush=True)
    assert info["fallback"] is True
    assert r[2] >= r[1]
    assert info["sweeps"] == len(r_plain) + 1
    np.testing.assert_allclose(np.r_[r[:2], r[3:]], r_plain, rtol=1e-6)
    np.testing.assert_allclose(x_fb, x_plain, rtol=0, atol=1e-6 * float(np.abs(x_plain).max()))
    assert r[-1] <= target


def test_later_stage_fallback_restores_the_kept_iterate(gpu_plan_factory):
    """A harmful second-stage update (3 sweeps after the first network's): its fallback is
    taken and the solve continues from the iterate before that update, i.e. it is the
    first-stage-only solve with one wasted sweep inserted."""
    from ctrecon_amd import geometry as G
    from ctrecon_amd import learned as L
    from ctrecon_amd.plan import DeviceArray, Plan, Sart
    from ctrecon_amd.unet import UNet

    print("[fallback, later stage] seed 30002, 256x256x128, bf16, target 2e-3", flush=True)
    g = G.dd3_geometry(G.config_training(), mask=G.fov_mask(256))
    vol = L.training_phantom(g, 30002)
    plan = Plan(g, 0)
    plan.set_storage("bf16")
    d_b = DeviceArray(g.prj_shape)
    plan.fp(DeviceArray.from_host(vol), d_b)
    target = 2e-3
    net = UNet.from_file(WEIGHTS)
    S = Sart(plan, d_b, 20)
    dx = DeviceArray(g.vol_shape)
    one = L.warm_start_os_sart(plan, net, d_b, dx, 20, target, 300, sart=S)
    x_one = dx.get()
    two = L.warm_start_os_sart(plan, net, d_b, dx, 20, target, 300, sart=S,
                               stages=[(_harmful(WEIGHTS), 3)])
    x_two = dx.get()
    r1, r2 = np.asarray(one["residuals"]), np.asarray(two["residuals"])
    print(f"  first stage only: {one['sweeps']} sweeps; with the harmful stage: {two['sweeps']} "
          f"sweeps, fallbacks {two['fallback']}/{two['stage_fallbacks']}", flush=True)
    assert len(r1) > 6, "the first stage alone must not reach the target before the stage acts"
    assert two["fallback"] is False and two["stage_fallbacks"] == [True]
    # sweeps 1-2, the update, 3 sweeps (the kept iterate), the wasted sweep, then the rest
    assert two["sweeps"] == one["sweeps"] + 1
    assert r2[5] >= r2[4]
    np.testing.assert_allclose(np.r_[r2[:5], r2[6:]], r1, rtol=1e-6)
    np.testing.assert_allclose(x_two, x_one, rtol=0, atol=1e-6 * float(np.abs(x_one).max()))

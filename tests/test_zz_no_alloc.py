"""ctr_residual_noalloc (the residual for callers that cannot allocate device memory)
against ctr_residual, through the C ABI: two calls in a row, the second on a smaller scan,
each equal to ctr_residual bit for bit (same view chunks, same kernels, same sum order);
with every device allocation of the library set to fail during the calls; and its errors
(integral-image method, a detector beyond the fixed scratch)."""

import ctypes

import numpy as np
import pytest

from cases import small, volume

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def _medium():
    from ctrecon_amd import geometry as G

    # more than 256 views (several chunks) and a 64-row detector (column-per-wave kernels)
    cfg = G.helical_config(sid=541, sdd=949, DNU=200, DNV=64, col_width=1.2, col_height=0.8,
                           views_per_turn=300, turns=2, feed_mm_per_turn=30.0, N_xy=96, N_z=48,
                           voxel_xy=0.9, voxel_z=0.9)
    return cfg, G.dd3_geometry(cfg, mask=G.fov_mask(96))


def _scan(make):
    from ctrecon_amd.plan import DeviceArray, Plan

    _, g = make()
    plan = Plan(g, 0)
    b = DeviceArray(g.prj_shape)
    plan.fp(DeviceArray.from_host(volume(g)), b)
    # an iterate some way from the answer: half the phantom plus noise
    rng = np.random.default_rng(3)
    x = (0.5 * volume(g) + rng.uniform(0, 0.05, g.vol_shape)).astype(np.float32)
    plan.synchronize()
    return g, plan, DeviceArray.from_host(x), b


def _noalloc(plan, d_x, d_b):
    from ctrecon_amd._lib import lib

    out = ctypes.c_double(-1.0)
    rc = lib().ctr_residual_noalloc(plan.handle, d_x.ptr, d_b.ptr, ctypes.byref(out))
    return rc, out.value


def test_noalloc_residual_equals_residual(gpu_plan_factory):
    from ctrecon_amd._lib import CTR_OK, lib
    from ctrecon_amd.plan import DeviceArray

    for name, make in (("medium 96^2x48, 600 views", _medium), ("small 64^3, 720 views", small)):
        g, plan, d_x, d_b = _scan(make)
        ref = plan.residual(d_x, d_b)
        # any device allocation from here on fails: the call must not make one
        lib().ctr_debug_fail_alloc_after(0)
        try:
            rc, val = _noalloc(plan, d_x, d_b)
            msg = lib().ctr_last_error().decode()
        finally:
            lib().ctr_debug_fail_alloc_after(-1)
        _say(f"[noalloc] {name}: rc {rc} ({msg!r}) residual {val!r} vs ctr_residual {ref!r}")
        assert rc == CTR_OK
        assert val == ref
        assert np.isfinite(val) and 0.0 < val < 1.0
        # again on the same GPU's scratch, after other work of the plan
        plan.fp(d_x, DeviceArray(g.prj_shape))
        rc2, val2 = _noalloc(plan, d_x, d_b)
        assert rc2 == CTR_OK and val2 == ref


def test_noalloc_residual_errors(gpu_plan_factory):
    from ctrecon_amd import geometry as G
    from ctrecon_amd._lib import CTR_ERR_ARG, CTR_ERR_UNSUPPORTED, lib
    from ctrecon_amd.plan import DeviceArray, Plan

    g, plan, d_x, d_b = _scan(small)
    plan.set_method("sat")
    rc, _ = _noalloc(plan, d_x, d_b)
    assert rc == CTR_ERR_UNSUPPORTED
    assert "direct method" in lib().ctr_last_error().decode()
    # 1024 x 256 rays per view and 256 views: beyond the 4 MiB scratch
    cfg = G.helical_config(sid=541, sdd=949, DNU=1024, DNV=256, col_width=0.5, col_height=0.2,
                           views_per_turn=256, turns=1, feed_mm_per_turn=10.0, N_xy=8, N_z=8,
                           voxel_xy=2.0, voxel_z=2.0)
    gb = G.dd3_geometry(cfg)
    big = Plan(gb, 0)
    x = DeviceArray(gb.vol_shape)
    x.zero()
    b = DeviceArray(gb.prj_shape)
    b.zero()
    rc, _ = _noalloc(big, x, b)
    msg = lib().ctr_last_error().decode()
    _say(f"[noalloc] oversized detector: rc {rc} {msg!r}")
    assert rc == CTR_ERR_ARG
    assert "4 MiB" in msg

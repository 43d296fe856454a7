"""Batched OS-SART (one persistent launch per subset step over several scans, sweeps
replayed from a HIP graph, tuned launch shapes) against per-scan OS-SART.

With the launch shapes of the per-scan path (one wave per detector column, one z group
per lane) every ray and voxel is computed by the same instruction sequence, so the
iterates must be bit-identical; tuned shapes regroup the z arithmetic of the back
projector and agree to FP32 rounding.  Graph replays must equal plain launches sweep
by sweep (the residual sums, whose partial slots are ordered differently, to 1e-12).
"""

import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu

SUBSETS = 5


def _say(msg):
    print(msg, flush=True)


def _scans():
    from ctrecon_amd import geometry as G
    from ctrecon_amd import phantom as P

    specs = [  # N_xy, N_z, voxel, views/turn, turns, feed/turn: uneven sizes and pitches
        (48, 32, 1.0, 90, 2, 20.0),
        (64, 48, 0.8, 120, 3, 26.0),
        (40, 64, 1.1, 100, 2, 16.0),
    ]
    out = []
    for k, (nxy, nz, vox, vpt, turns, feed) in enumerate(specs):
        cfg = G.helical_config(sid=541, sdd=949, DNU=96, DNV=64, col_width=1.5, col_height=1.0,
                               views_per_turn=vpt, turns=turns, feed_mm_per_turn=feed, N_xy=nxy,
                               N_z=nz, voxel_xy=vox, voxel_z=vox)
        g = G.dd3_geometry(cfg, mask=G.fov_mask(nxy))
        vol = P.shepp_logan(g, seed=k)
        out.append((g, vol))
    return out


def _setup(storage="f32"):
    from ctrecon_amd.plan import DeviceArray, Plan

    plans, xs, bs = [], [], []
    for g, vol in _scans():
        p = Plan(g, 0)
        p.set_storage(storage)
        b = DeviceArray(g.prj_shape)
        p.fp(DeviceArray.from_host(vol), b)
        x = DeviceArray(g.vol_shape)
        x.zero()
        plans.append(p)
        xs.append(x)
        bs.append(b)
    return plans, xs, bs


def _per_scan(plans, bs, sweeps):
    from ctrecon_amd.plan import DeviceArray

    out = []
    for p, b in zip(plans, bs):
        x = DeviceArray(p.g.vol_shape)
        x.zero()
        r = p.os_sart(x, b, SUBSETS, sweeps)
        out.append((x.get(), r))
    return out


@pytest.mark.parametrize("storage", ["f32", "bf16"])
def test_batch_equals_per_scan(storage, gpu_plan_factory):
    from ctrecon_amd.batch import Batch

    _say(f"[batch = per-scan] storage={storage}")
    plans, xs, bs = _setup(storage)
    ref = _per_scan(plans, bs, 3)
    B = Batch(plans, xs, bs, SUBSETS)
    # the per-scan launch shapes: z-group back projector (all ZN <= 64), then the
    # windowed one; both compute every voxel's sum with the same arithmetic in view order
    for zg in (1, 5):
        for x in xs:
            x.zero()
        B.set_tuning(True, 512, zg, 512)
        res, ms = B.sweeps(3)
        for i, (x_ref, r_ref) in enumerate(ref):
            x = xs[i].get()
            _say(f"  zg {zg} scan {i}: max |batch - per-scan| = "
                 f"{np.max(np.abs(x - x_ref)):.3e}; residuals {res[:, i]} vs {r_ref}")
            assert np.array_equal(x, x_ref)
            # the FP64 residual partials are summed in another slot order
            assert np.allclose(res[:, i], r_ref, rtol=1e-12, atol=0)


def test_graph_replay_equals_launches(gpu_plan_factory):
    from ctrecon_amd.batch import Batch

    _say("[batch graph replay]")
    plans, xs, bs = _setup()
    B = Batch(plans, xs, bs, SUBSETS)
    res_g, _ = B.sweeps(4)
    xg = [x.get() for x in xs]
    for x in xs:
        x.zero()
    B.set_graph(False)
    res_e, _ = B.sweeps(4)
    _say(f"  info {B.info()}")
    _say(f"  residuals graph {res_g[:, 0]} launches {res_e[:, 0]}")
    assert np.array_equal(res_g, res_e)
    for i, x in enumerate(xs):
        assert np.array_equal(x.get(), xg[i])


def test_tuned_batch_matches_per_scan(gpu_plan_factory):
    from ctrecon_amd.batch import Batch

    _say("[tuned batch]")
    plans, xs, bs = _setup()
    ref = _per_scan(plans, bs, 3)
    B = Batch(plans, xs, bs, SUBSETS)
    info = B.info()
    res, ms = B.sweeps(3)
    _say(f"  tuning {info}; sweep ms {ms}")
    for i, (x_ref, r_ref) in enumerate(ref):
        e = rel_l2(xs[i].get(), x_ref)
        _say(f"  scan {i}: rel-L2 vs per-scan {e:.3e}")
        assert e <= 1e-5
        assert np.allclose(res[:, i], r_ref, rtol=1e-5)
    # every tuning candidate computes the same iterates
    for colw, zg, grid, strided in [(True, 0, 0, False), (True, 1, 1024, False),
                                    (False, 2, 2048, True),
                                    (True, 4, 0, False), (False, 1, 0, True),
                                    (True, 2, 512, True), (True, 5, 0, False),
                                    (False, 5, 1024, True), (True, 5, 512, False)]:
        for x in xs:
            x.zero()
        B.set_tuning(colw, grid, zg, grid, strided, strided)
        B.sweeps(3)
        for i, (x_ref, _) in enumerate(ref):
            assert rel_l2(xs[i].get(), x_ref) <= 1e-5, (colw, zg, grid, strided, i)

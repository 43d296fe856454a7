"""Solver plumbing on the GPU: sessions replay captured sweep graphs, CG keeps its vectors
in the plan, device-allocation failures surface as CTR_ERR_NOMEM and leave the library
usable, and the batched autotuner's choice is cached per geometry."""

import ctypes

import numpy as np
import pytest

from cases import small, volume

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def _problem(storage="f32"):
    from ctrecon_amd.plan import DeviceArray, Plan

    cfg, g = small()
    plan = Plan(g, 0)
    plan.set_storage(storage)
    b = DeviceArray(g.prj_shape)
    plan.fp(DeviceArray.from_host(volume(g)), b)
    return g, plan, b


@pytest.mark.parametrize("storage", ["f32", "bf16"])
def test_graph_replay_equals_plain_sweeps(storage, gpu_plan_factory):
    """Sweeps replayed from the captured graph equal plain launches bit for bit, replay
    after replay, with other work of the program between the calls."""
    from ctrecon_amd.plan import DeviceArray, Sart

    _say(f"[sweep graph] small case, storage {storage}")
    g, plan, b = _problem(storage)
    S = Sart(plan, b, 10)
    x = DeviceArray(g.vol_shape)
    x.zero()
    res_g = []
    xs_g = []
    for k in range(4):
        r, ms, _ = S.sweeps(x, 1 + (k % 2))
        res_g += list(r)
        xs_g.append(x.get())
        # eager work of the program between replays
        tmp = DeviceArray(g.prj_shape)
        plan.fp(x, tmp)
        plan.synchronize()
    captured, cap_ms = S.graph_info()
    _say(f"  graph captured {captured} ({cap_ms:.1f} ms), residuals {res_g}")
    assert captured
    S.set_graph(False)
    x.zero()
    res_e, xs_e = [], []
    for k in range(4):
        r, _, _ = S.sweeps(x, 1 + (k % 2))
        res_e += list(r)
        xs_e.append(x.get())
    assert not S.graph_info()[0]
    assert res_g == res_e
    for a, c in zip(xs_g, xs_e):
        assert np.array_equal(a, c)
    # with per-launch timing the sweeps run as plain launches and still agree
    S.set_graph(True)
    x.zero()
    r, _, st = S.sweeps(x, 3, stats=True)
    assert st["fp_launches"] == 30 and list(r) == res_e[:3]


def test_new_iterate_is_captured_again(gpu_plan_factory):
    from ctrecon_amd.plan import DeviceArray, Sart

    g, plan, b = _problem()
    S = Sart(plan, b, 10)
    x1, x2 = DeviceArray(g.vol_shape), DeviceArray(g.vol_shape)
    x1.zero()
    x2.zero()
    r1, _, _ = S.sweeps(x1, 2)
    r2, _, _ = S.sweeps(x2, 2)  # another buffer: a new capture
    assert list(r1) == list(r2)
    assert np.array_equal(x1.get(), x2.get())


def test_conjgrad_reuses_its_buffers(gpu_plan_factory):
    """The second CG call on a plan allocates nothing (an injected allocation failure is
    not reached) and repeats the first call's result."""
    from ctrecon_amd._lib import lib
    from ctrecon_amd.plan import DeviceArray

    _say("[CG buffers] small case, 3 iterations, twice")
    g, plan, b = _problem()
    x = DeviceArray(g.vol_shape)
    x.zero()
    r1 = plan.conjgrad(x, b, 3)
    x1 = x.get()
    x.zero()
    lib().ctr_debug_fail_alloc_after(0)
    try:
        r2 = plan.conjgrad(x, b, 3)
    finally:
        lib().ctr_debug_fail_alloc_after(-1)
    assert list(r1) == list(r2)
    assert np.array_equal(x.get(), x1)


def test_allocation_failures_are_reported_and_recovered(gpu_plan_factory):
    """Every device allocation of plan creation and of an OS-SART session, failed in turn:
    the call returns CTR_ERR_NOMEM with a message, and afterwards the same calls succeed
    and compute what they compute without the injection."""
    from ctrecon_amd._lib import CTR_ERR_NOMEM, CtrError, lib
    from ctrecon_amd.plan import DeviceArray, Plan, Sart

    g, plan, b = _problem()
    x = DeviceArray(g.vol_shape)
    x.zero()
    ref = plan.os_sart(x, b, 10, 2)
    x_ref = x.get()
    failures = {"plan": 0, "session": 0}
    for what in ("plan", "session"):
        for k in range(40):
            lib().ctr_debug_fail_alloc_after(k)
            try:
                if what == "plan":
                    Plan(g, 0).close()
                else:
                    Sart(plan, b, 10).close()
            except CtrError as e:
                assert e.code == CTR_ERR_NOMEM, str(e)
                assert "allocat" in str(e)
                failures[what] += 1
                continue
            finally:
                lib().ctr_debug_fail_alloc_after(-1)
            break  # k is past the call's last allocation: it succeeded
        else:
            pytest.fail(f"{what}: more than 40 allocations")
    _say(f"  allocations failed in turn: {failures}")
    assert failures["plan"] >= 10 and failures["session"] >= 4
    x.zero()
    again = plan.os_sart(x, b, 10, 2)
    assert list(again) == list(ref)
    assert np.array_equal(x.get(), x_ref)
    # the host-buffer entry point fails cleanly too
    from ctrecon_amd import dd3 as D

    argv, keep = D._geo_argv(g)
    m = np.ascontiguousarray(g.mask, dtype=np.uint8)
    vol = np.ascontiguousarray(volume(g).ravel())
    prj = np.empty(g.PN * g.DNU * g.DNV, np.float32)
    lib().ctr_debug_fail_alloc_after(0)
    try:
        rc = lib().ctr_dd3_proj(*argv, vol.ctypes.data, prj.ctypes.data, g.dx, g.dz,
                                m.ctypes.data, 0, 0)
    finally:
        lib().ctr_debug_fail_alloc_after(-1)
    assert rc == CTR_ERR_NOMEM
    assert lib().ctr_dd3_proj(*argv, vol.ctypes.data, prj.ctypes.data, g.dx, g.dz,
                              m.ctypes.data, 0, 0) == 0


def test_batch_tuning_is_cached(tmp_path, monkeypatch, gpu_plan_factory):
    """The second batch of the same scans takes its launch shapes from the tuning cache
    (no timing) and computes the same sweeps."""
    from ctrecon_amd import geometry as G
    from ctrecon_amd import phantom as P
    from ctrecon_amd.batch import Batch
    from ctrecon_amd.plan import DeviceArray, Plan

    cache = tmp_path / "tuning.txt"
    monkeypatch.setenv("CTRECON_TUNE_CACHE", str(cache))
    plans, xs, bs = [], [], []
    for k, (nxy, nz) in enumerate([(48, 32), (40, 64)]):
        cfg = G.helical_config(sid=541, sdd=949, DNU=96, DNV=64, col_width=1.5, col_height=1.0,
                               views_per_turn=90, turns=2, feed_mm_per_turn=20.0, N_xy=nxy,
                               N_z=nz, voxel_xy=1.0, voxel_z=1.0)
        g = G.dd3_geometry(cfg, mask=G.fov_mask(nxy))
        p = Plan(g, 0)
        b = DeviceArray(g.prj_shape)
        p.fp(DeviceArray.from_host(P.shepp_logan(g, seed=k)), b)
        x = DeviceArray(g.vol_shape)
        x.zero()
        plans.append(p)
        xs.append(x)
        bs.append(b)
    B1 = Batch(plans, xs, bs, 5)
    i1 = B1.info()
    r1, _ = B1.sweeps(2)
    x1 = [x.get() for x in xs]
    B1.close()
    assert i1["from_cache"] == 0
    assert cache.exists() and len(cache.read_text().splitlines()) == 1
    for x in xs:
        x.zero()
    B2 = Batch(plans, xs, bs, 5)
    i2 = B2.info()
    _say(f"  tuned {i1}\n  cached {i2}")
    assert i2["from_cache"] == 1
    for k in ("fp_colw", "fp_grid", "fp_strided", "bp_zg", "bp_grid", "bp_strided"):
        assert i1[k] == i2[k]
    r2, _ = B2.sweeps(2)
    assert np.array_equal(r1, r2)
    for a, x in zip(x1, xs):
        assert np.array_equal(a, x.get())

"""GPU edge cases of the projector pair: columns taller than the windowed back
projectors' LDS limit (the z-group kernels then run), launches over zero views, and a
single-view scan, each against the FP64 oracle or the defined empty result."""

import numpy as np
import pytest

from conftest import rel_l2
from cases import sinogram, volume

from ctrecon_amd import geometry as G

pytestmark = pytest.mark.gpu

TOL = 1e-5


def _say(msg):
    print(msg, flush=True)


def _tall():
    """4 x 4 x 2100 voxels (0.08 mm slices): more than 2048 voxels per column."""
    cfg = G.helical_config(sid=300, sdd=520, DNU=8, DNV=4, col_width=2.0, col_height=2.0,
                           views_per_turn=12, turns=4, feed_mm_per_turn=40.0, N_xy=4, N_z=2100,
                           voxel_xy=4.0, voxel_z=0.08)
    return cfg, G.dd3_geometry(cfg)


def test_tall_columns_use_the_group_kernel(oracle, gpu_plan_factory):
    _say("[tall columns] 4x4x2100")
    cfg, g = _tall()
    plan = gpu_plan_factory(g)
    rng = np.random.default_rng(11)
    vol = rng.uniform(0, 1, g.vol_shape).astype(np.float32)
    y = sinogram(g)
    ef = rel_l2(plan.forward(vol), oracle.fp(g, vol, "sat"))
    back = plan.back(y, kind="ref")
    ref = oracle.bp(g, y, "sat")
    eb = rel_l2(back, ref)
    refm = oracle.bp(g, y, "adjoint")
    em = rel_l2(plan.back(y, kind="matched"), refm)
    _say(f"  fp rel-L2 {ef:.3e}, bp-ref rel-L2 {eb:.3e}, matched rel-L2 {em:.3e}")
    assert np.any(ref != 0) and np.any(refm != 0)
    assert ef <= TOL and eb <= TOL and em <= TOL


def test_zero_view_launches(gpu_plan_factory):
    """fp over no views writes nothing; bp over no views is the empty sum (0)."""
    from cases import tiny
    from ctrecon_amd.plan import DeviceArray

    cfg, g = tiny()
    plan = gpu_plan_factory(g)
    d_vol = DeviceArray.from_host(volume(g))
    sentinel = np.full((2, g.DNU, g.DNV), 7.0, np.float32)
    d_prj = DeviceArray.from_host(sentinel)
    plan.fp(d_vol, d_prj, 0, 1, 0)
    plan.synchronize()
    assert np.array_equal(d_prj.get(), sentinel)
    d_out = DeviceArray.from_host(np.full(g.vol_shape, 3.0, np.float32))
    plan.bp(d_prj, d_out, 0, 1, 0, "ref")
    plan.synchronize()
    assert np.all(d_out.get() == 0)


@pytest.mark.parametrize("p", [0, 5, 13])
def test_single_view_scan(p, oracle, gpu_plan_factory):
    """A scan of one view (the plan's tables sized for PN = 1)."""
    from cases import ragged

    cfg, g = ragged()
    g1 = g.subset([p])
    plan = gpu_plan_factory(g1)
    vol = volume(g1)
    y = sinogram(g1)
    ef = rel_l2(plan.forward(vol), oracle.fp(g1, vol, "sat"))
    eb = rel_l2(plan.back(y, kind="ref"), oracle.bp(g1, y, "sat"))
    _say(f"[single view {p}] fp {ef:.3e} bp {eb:.3e}")
    assert ef <= TOL and eb <= TOL


@pytest.mark.parametrize("col_width,col_height,DNU,DNV",
                         [(4.0, 0.8, 15, 15), (0.5, 0.4, 120, 30), (0.5, 3.0, 120, 4)])
def test_matched_many_cells_per_voxel(col_width, col_height, DNU, DNV, oracle,
                                      gpu_plan_factory):
    """Voxels spanning many detector cells: more rows per voxel than the matched
    adjoint's fixed-count row gather takes (its row walk runs), and more than four
    columns per voxel (its recomputed-weight path); against the FP64 oracle adjoint and
    through the adjoint identity."""
    _say(f"[matched many cells] col {col_width} x {col_height} mm, {DNU} x {DNV}")
    cfg = G.helical_config(sid=60, sdd=100, DNU=DNU, DNV=DNV, col_width=col_width,
                           col_height=col_height, views_per_turn=8, turns=2,
                           feed_mm_per_turn=6.0, N_xy=8, N_z=8, voxel_xy=4.0, voxel_z=3.0)
    g = G.dd3_geometry(cfg)
    plan = gpu_plan_factory(g)
    y = sinogram(g)
    got = plan.back(y, kind="matched")
    ref = oracle.bp(g, y, "adjoint")
    err = rel_l2(got, ref)
    rng = np.random.default_rng(5)
    x = rng.uniform(0, 1, g.vol_shape).astype(np.float32)
    lhs = float(np.dot(plan.forward(x).astype(np.float64).ravel(), y.astype(np.float64).ravel()))
    rhs = float(np.dot(x.astype(np.float64).ravel(), got.astype(np.float64).ravel()))
    adj = abs(lhs - rhs) / abs(lhs)
    _say(f"  matched rel-L2 {err:.3e}, adjoint identity {adj:.3e}")
    assert np.any(ref != 0)
    assert err <= TOL and adj <= 1e-6

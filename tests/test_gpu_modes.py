"""GPU parity of the reference's other projector modes against the FP64 oracle: the
pseudo distance-driven pair (prjMode 1 / 3 forward, 1 / 2 back), the FP64 branchless
forward projector (prjMode 2) and the z-line back projector (prjMode 3), through the
plan API and the DD3Proj_gpu / DD3Back_gpu entry points.  Bar: 1e-5 relative L2 (FP32
arithmetic against FP64).
"""

import numpy as np
import pytest

from conftest import rel_l2
from cases import CASES, clinical, sinogram, small, tiny, volume

pytestmark = pytest.mark.gpu

TOL = 1e-5


def _say(msg):
    print(msg, flush=True)


def _dd3_proj(g, vol, mode):
    from ctrecon_amd import dd3 as D
    from ctrecon_amd._lib import check, lib

    argv, keep = D._geo_argv(g)
    m = np.ascontiguousarray(g.mask, dtype=np.uint8)
    v = np.ascontiguousarray(vol, dtype=np.float32)
    out = np.empty(g.PN * g.DNU * g.DNV, np.float32)
    check(lib().ctr_dd3_proj(*argv, v.ctypes.data, out.ctypes.data, g.dx, g.dz, m.ctypes.data,
                             0, mode))
    return out.reshape(g.prj_shape)


def _dd3_back(g, prj, mode):
    from ctrecon_amd import dd3 as D
    from ctrecon_amd._lib import check, lib

    argv, keep = D._geo_argv(g)
    m = np.ascontiguousarray(g.mask, dtype=np.uint8)
    y = np.ascontiguousarray(prj, dtype=np.float32)
    out = np.empty(g.XN * g.YN * g.ZN, np.float32)
    check(lib().ctr_dd3_back(*argv, out.ctypes.data, y.ctypes.data, g.dx, g.dz, m.ctypes.data,
                             0, 0, mode))
    return out.reshape(g.vol_shape)


@pytest.mark.parametrize("name", list(CASES))
def test_pdd_fp_parity(name, oracle, gpu_plan_factory):
    _say(f"[pdd fp] case={name}")
    cfg, g = CASES[name]()
    if name == "small":
        g = g.subset(range(0, g.PN, 9))
    plan = gpu_plan_factory(g)
    plan.set_method("pseudo_dd")
    vol = volume(g)
    got = plan.forward(vol)
    ref = oracle.fp(g, vol, "pdd")
    err = rel_l2(got, ref)
    _say(f"  pdd fp rel-L2 vs FP64 oracle = {err:.3e}")
    assert err <= TOL


@pytest.mark.parametrize("name", list(CASES))
def test_pdd_bp_parity(name, oracle, gpu_plan_factory):
    _say(f"[pdd bp] case={name}")
    cfg, g = CASES[name]()
    if name == "small":
        g = g.subset(range(0, g.PN, 9))
    plan = gpu_plan_factory(g)
    plan.set_method("pseudo_dd")
    y = sinogram(g)
    got = plan.back(y, kind="ref")
    ref = oracle.bp(g, y, "pdd")
    err = rel_l2(got, ref)
    _say(f"  pdd bp rel-L2 vs FP64 oracle = {err:.3e}")
    assert err <= TOL
    m = g.mask.reshape(g.YN, g.XN)
    assert np.all(got[m != 1] == 0)


@pytest.mark.parametrize("name", ["tiny", "ragged", "flat"])
def test_reference_entry_points_select_modes(name, oracle):
    """DD3Proj_gpu / DD3Back_gpu prjMode dispatch (DD3_GPU.cu:1144-1183, 1778-1815)."""
    _say(f"[prjMode dispatch] case={name}")
    cfg, g = CASES[name]()
    vol = volume(g)
    y = sinogram(g)
    f_pdd = oracle.fp(g, vol, "pdd")
    for mode in (1, 3):
        assert rel_l2(_dd3_proj(g, vol, mode), f_pdd) <= TOL
    err = rel_l2(_dd3_proj(g, vol, 2), oracle.fp(g, vol, "dp"))
    _say(f"  prjMode 2 forward rel-L2 vs FP64 oracle = {err:.3e}")
    assert err <= TOL
    b_pdd = oracle.bp(g, y, "pdd")
    for mode in (1, 2):
        assert rel_l2(_dd3_back(g, y, mode), b_pdd) <= TOL
    got = _dd3_back(g, y, 3)
    ref = oracle.bp(g, y, "zline")
    if np.any(ref != 0):
        err = rel_l2(got, ref)
        _say(f"  prjMode 3 back rel-L2 vs FP64 oracle = {err:.3e}")
        assert err <= TOL
    else:
        assert np.all(got == 0)
    # values outside 1..3 are mode 0
    assert rel_l2(_dd3_proj(g, vol, 7), oracle.fp(g, vol, "sat")) <= TOL


def test_fp64_sat_mode_small(oracle):
    cfg, g = small()
    g = g.subset(range(0, g.PN, 13))
    vol = volume(g)
    err = rel_l2(_dd3_proj(g, vol, 2), oracle.fp(g, vol, "dp"))
    _say(f"[prjMode 2 small] rel-L2 = {err:.3e}")
    assert err <= TOL


def test_pdd_clinical_sampled_views(oracle, gpu_plan_factory):
    """Clinical geometry (512^2 x 256, 888 x 64): pseudo-DD forward projection of one view
    per marching direction and the back projection of one view."""
    from ctrecon_amd import phantom as P

    _say("[pdd clinical sampled views] building phantom")
    cfg, g = clinical()
    vol = P.ellipsoid_mixture(g, seed=1)
    plan = gpu_plan_factory(g)
    plan.set_method("pseudo_dd")
    views = [0, 246]
    xm = [abs(np.cos(g.angles[p])) <= abs(np.sin(g.angles[p])) for p in views]
    assert xm[0] != xm[1]
    for p in views:
        got = plan.forward(vol, [p])
        err = rel_l2(got, oracle.fp(g.subset([p]), vol, "pdd"))
        _say(f"  view {p}: pdd fp rel-L2 = {err:.3e}")
        assert err <= TOL
    y = sinogram(g.subset([0]))
    got = plan.back(y, [0], kind="ref")
    err = rel_l2(got, oracle.bp(g.subset([0]), y, "pdd"))
    _say(f"  pdd bp one view rel-L2 = {err:.3e}")
    assert err <= TOL


def test_pdd_plan_refusals(gpu_plan_factory):
    """No matched adjoint, no BF16 operands, no OS-SART / CG with the pseudo-DD method."""
    from ctrecon_amd._lib import CtrError
    from ctrecon_amd.plan import DeviceArray

    cfg, g = tiny()
    plan = gpu_plan_factory(g)
    plan.set_method("pseudo_dd")
    y = sinogram(g)
    with pytest.raises(CtrError):
        plan.back(y, kind="matched")
    with pytest.raises(CtrError):
        plan.forward_bf16(volume(g))
    d_x = DeviceArray.from_host(np.zeros(g.vol_shape, np.float32))
    d_b = DeviceArray.from_host(y)
    with pytest.raises(CtrError):
        plan.os_sart(d_x, d_b, 2, 1)
    with pytest.raises(CtrError):
        plan.conjgrad(d_x, d_b, 1)
    plan.set_method("direct")
    plan.os_sart(d_x, d_b, 2, 1)


def _tall(ZN):
    from ctrecon_amd import geometry as G

    cfg = G.helical_config(sid=300, sdd=520, DNU=40, DNV=8, col_width=2.0, col_height=1.6,
                           views_per_turn=36, turns=2, feed_mm_per_turn=8.0, N_xy=24, N_z=ZN,
                           voxel_xy=2.0, voxel_z=1.0)
    m = np.ones((24, 24), np.uint8)
    m[np.random.default_rng(2).uniform(size=m.shape) < 0.1] = 0
    return cfg, G.dd3_geometry(cfg, mask=m)


@pytest.mark.parametrize("ZN", [37, 64, 150])
def test_zline_parity(ZN, oracle, gpu_plan_factory):
    """ZN 37 (two whole 16-slice blocks, 5 slices left 0), 64 (one z group), 150 (three
    z groups in one wave, 6 slices left 0)."""
    _say(f"[z-line] ZN={ZN}")
    cfg, g = _tall(ZN)
    plan = gpu_plan_factory(g)
    plan.set_method("zline")
    y = sinogram(g)
    got = plan.back(y, kind="ref")
    ref = oracle.bp(g, y, "zline")
    err = rel_l2(got, ref)
    _say(f"  z-line rel-L2 vs FP64 oracle = {err:.3e}")
    assert err <= TOL
    NZ = 16 * (ZN // 16)
    assert np.all(got[:, :, NZ:] == 0)
    m = g.mask.reshape(g.YN, g.XN)
    assert np.all(got[m != 1] == 0)
    # the prjMode 3 forward projector is the pseudo-DD one
    vol = volume(g)
    assert rel_l2(plan.forward(vol), oracle.fp(g, vol, "pdd")) <= TOL


def test_zline_clinical_view(oracle, gpu_plan_factory):
    """Clinical geometry (512^2 x 256, 888 x 64): z-line back projection of one view."""
    cfg, g = clinical()
    plan = gpu_plan_factory(g)
    plan.set_method("zline")
    y = sinogram(g.subset([0]))
    got = plan.back(y, [0], kind="ref")
    err = rel_l2(got, oracle.bp(g.subset([0]), y, "zline"))
    _say(f"[z-line clinical view 0] rel-L2 = {err:.3e}")
    assert err <= TOL

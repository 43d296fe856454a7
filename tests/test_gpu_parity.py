"""GPU parity of the HIP projector pair and loops against the FP64 oracle.

Bar (from the task): forward and back projections within 1e-5 relative L2 of
the FP64 oracle in FP32, and <Ax, y> = <x, A^T y> to 1e-6 with FP64 dots.
"""

import numpy as np
import pytest

from conftest import rel_l2
from cases import CASES, clinical, sinogram, small, tiny, volume

pytestmark = pytest.mark.gpu

FP_TOL = 1e-5
BP_TOL = 1e-5
ADJ_TOL = 1e-6


def _say(msg):
    print(msg, flush=True)


@pytest.mark.parametrize("name", list(CASES))
def test_fp_parity(name, oracle, gpu_plan_factory):
    _say(f"[fp parity] case={name}")
    cfg, g = CASES[name]()
    plan = gpu_plan_factory(g)
    vol = volume(g)
    got = plan.forward(vol)
    ref = oracle.fp(g, vol, "sat")
    err = rel_l2(got, ref)
    _say(f"  fp rel-L2 vs FP64 oracle = {err:.3e}")
    assert err <= FP_TOL


@pytest.mark.parametrize("name", list(CASES))
def test_bp_reference_parity(name, oracle, gpu_plan_factory):
    _say(f"[bp-ref parity] case={name}")
    cfg, g = CASES[name]()
    plan = gpu_plan_factory(g)
    y = sinogram(g)
    got = plan.back(y, kind="ref")
    ref = oracle.bp(g, y, "sat")
    err = rel_l2(got, ref)
    _say(f"  bp-ref rel-L2 vs FP64 oracle = {err:.3e}")
    assert err <= BP_TOL
    m = g.mask.reshape(g.YN, g.XN)
    assert np.all(got[m != 1] == 0)


@pytest.mark.parametrize("name", list(CASES))
def test_bp_matched_parity(name, oracle, gpu_plan_factory):
    _say(f"[bp-matched parity] case={name}")
    cfg, g = CASES[name]()
    plan = gpu_plan_factory(g)
    y = sinogram(g)
    got = plan.back(y, kind="matched")
    ref = oracle.bp(g, y, "adjoint")
    err = rel_l2(got, ref)
    _say(f"  matched rel-L2 vs FP64 oracle adjoint = {err:.3e}")
    assert err <= BP_TOL


@pytest.mark.parametrize("name", list(CASES))
def test_adjoint_identity(name, gpu_plan_factory):
    _say(f"[adjoint identity] case={name}")
    cfg, g = CASES[name]()
    plan = gpu_plan_factory(g)
    rng = np.random.default_rng(3)
    x = rng.uniform(0, 1, g.vol_shape).astype(np.float32)
    y = rng.uniform(0, 1, g.prj_shape).astype(np.float32)
    Ax = plan.forward(x).astype(np.float64)
    Aty = plan.back(y, kind="matched").astype(np.float64)
    lhs = float(np.dot(Ax.ravel(), y.astype(np.float64).ravel()))
    rhs = float(np.dot(x.astype(np.float64).ravel(), Aty.ravel()))
    err = abs(lhs - rhs) / abs(lhs)
    _say(f"  <Ax,y>={lhs:.10e} <x,A^T y>={rhs:.10e} rel={err:.3e}")
    assert err <= ADJ_TOL


def test_subset_views(oracle, gpu_plan_factory):
    """Strided view subsets (the OS-SART access pattern) project like the full scan."""
    cfg, g = small()
    plan = gpu_plan_factory(g)
    vol = volume(g)
    views = range(3, g.PN, 10)
    got = plan.forward(vol, views)
    ref = oracle.fp(g.subset(list(views)), vol, "sat")
    assert rel_l2(got, ref) <= FP_TOL
    y = sinogram(g.subset(list(views)))
    got_b = plan.back(y, views, kind="ref")
    ref_b = oracle.bp(g.subset(list(views)), y, "sat")
    assert rel_l2(got_b, ref_b) <= BP_TOL


@pytest.mark.parametrize("kind", ["ref", "matched"])
def test_os_sart_parity(kind, oracle, gpu_plan_factory):
    from ctrecon_amd.plan import DeviceArray

    _say(f"[os-sart parity] bp={kind}")
    cfg, g = small()
    plan = gpu_plan_factory(g)
    vol = volume(g)
    b = plan.forward(vol)
    dx = DeviceArray(g.vol_shape)
    dx.zero()
    db = DeviceArray.from_host(b)
    res = plan.os_sart(dx, db, 10, 2, kind)
    x = dx.get()
    xr, res_r = oracle.os_sart(g, b, osNum=10, iterNum=2, bp_kind=kind)
    err = rel_l2(x, xr)
    _say(f"  iterate rel-L2 vs FP64 oracle = {err:.3e}; residuals gpu {res} oracle {res_r}")
    assert err <= 1e-4
    assert np.allclose(res, res_r, rtol=1e-3)


@pytest.mark.parametrize("name", ["ragged", "flat"])
def test_os_sart_first_sweep_parity(name, oracle, gpu_plan_factory):
    """One sweep from x = 0: the row and column sums (evaluated on the GPU from the
    footprints' closed-form sums of ones, mask bytes 0 / 1 / 2 in the ragged case) and
    the first update against the FP64 oracle, which projects arrays of ones."""
    from ctrecon_amd.plan import DeviceArray

    _say(f"[os-sart first sweep] case={name}")
    cfg, g = CASES[name]()
    plan = gpu_plan_factory(g)
    b = plan.forward(volume(g))
    dx = DeviceArray(g.vol_shape)
    dx.zero()
    res = plan.os_sart(dx, DeviceArray.from_host(b), 5, 1, "ref")
    xr, res_r = oracle.os_sart(g, b, osNum=5, iterNum=1, bp_kind="ref")
    err = rel_l2(dx.get(), xr)
    _say(f"  iterate rel-L2 vs FP64 oracle = {err:.3e}; residual gpu {res} oracle {res_r}")
    assert err <= 1e-5
    assert np.allclose(res, res_r, rtol=1e-5)


@pytest.mark.parametrize("kind", ["ref", "matched"])
def test_conjgrad_parity(kind, oracle, gpu_plan_factory):
    from ctrecon_amd.plan import DeviceArray

    _say(f"[cg parity] bp={kind}")
    cfg, g = small()
    plan = gpu_plan_factory(g)
    vol = volume(g)
    b = plan.forward(vol)
    dx = DeviceArray(g.vol_shape)
    dx.zero()
    db = DeviceArray.from_host(b)
    res = plan.conjgrad(dx, db, 3, kind)
    x = dx.get()
    xr, res_r = oracle.conjgrad(g, b, iterNum=3, bp_kind=kind)
    err = rel_l2(x, xr)
    _say(f"  iterate rel-L2 vs FP64 oracle = {err:.3e}; |r| gpu {res} oracle {res_r}")
    assert err <= 1e-4


def test_dd3_matlab_surface(gpu_plan_factory):
    """dd3('fp_gpu_branchless' | 'bp_gpu_branchless', ...) through the reference-signature
    entry points ctr_dd3_proj / ctr_dd3_back equals the plan path."""
    from ctrecon_amd import dd3 as D

    cfg, g = small()
    plan = gpu_plan_factory(g)
    vol = volume(g)  # GPU layout [y, x, z]
    img = np.transpose(vol, (2, 0, 1))  # MATLAB [Z, X, Y]: GPU y is MATLAB x (dd3.m:248)
    m = g.mask.reshape(g.YN, g.XN).T
    sino = D.dd3("fp_gpu_branchless", cfg, img, None, m)
    assert sino.shape == (g.DNV, g.DNU, g.PN)
    ref = plan.forward(vol)
    assert rel_l2(np.transpose(sino, (2, 1, 0)), ref) <= 1e-6
    back = D.dd3("bp_gpu_branchless", cfg, sino, None, m)
    assert back.shape == img.shape
    refb = plan.back(np.ascontiguousarray(np.transpose(sino, (2, 1, 0))), kind="ref")
    assert rel_l2(np.transpose(back, (1, 2, 0)), refb) <= 1e-6


def test_reference_mode_errors(gpu_plan_factory):
    """The matched adjoint (bit 4) exists for the branchless model only: prjMode 1 | 16
    is refused; 0 | 16 runs."""
    from ctrecon_amd import dd3 as D
    from ctrecon_amd._lib import CtrError, lib

    cfg, g = tiny()
    argv, keep = D._geo_argv(g)
    m = np.ascontiguousarray(g.mask, dtype=np.uint8)
    vol = np.empty(g.XN * g.YN * g.ZN, np.float32)
    prj = np.ones(g.PN * g.DNU * g.DNV, np.float32)
    rc = lib().ctr_dd3_back(*argv, vol.ctypes.data, prj.ctypes.data, g.dx, g.dz, m.ctypes.data,
                            0, 0, 1 | 16)
    assert rc == 3
    rc = lib().ctr_dd3_back(*argv, vol.ctypes.data, prj.ctypes.data, g.dx, g.dz, m.ctypes.data,
                            0, 0, 0 | 16)
    assert rc == 0


def test_clinical_sampled_views(oracle, gpu_plan_factory):
    """Clinical geometry (512^2 x 256, 888 x 64): forward projection of two views (one per
    marching direction) and the back projection of those views, against the oracle."""
    from ctrecon_amd import phantom as P

    _say("[clinical sampled views] building phantom")
    cfg, g = clinical()
    vol = P.ellipsoid_mixture(g, seed=1)
    plan = gpu_plan_factory(g)
    xm = [p for p in (0, 123, 246, 500) if abs(np.cos(g.angles[p])) <= abs(np.sin(g.angles[p]))]
    ym = [p for p in (0, 123, 246, 500) if abs(np.cos(g.angles[p])) > abs(np.sin(g.angles[p]))]
    for p in (xm[0], ym[0]):
        got = plan.forward(vol, [p])
        ref = oracle.fp(g.subset([p]), vol, "direct")
        err = rel_l2(got, ref)
        _say(f"  view {p}: fp rel-L2 = {err:.3e}")
        assert err <= FP_TOL
    y = sinogram(g.subset([xm[0]]))
    got = plan.back(y, [xm[0]], kind="ref")
    ref = oracle.bp(g.subset([xm[0]]), y, "direct")
    err = rel_l2(got, ref)
    _say(f"  bp-ref one view rel-L2 = {err:.3e}")
    assert err <= BP_TOL


# ---------------------------------------------------------------- integral-image method
@pytest.mark.parametrize("name", list(CASES))
def test_sat_method_parity(name, oracle, gpu_plan_factory):
    """The integral-image method (compensated FP32 SATs built with wave scans, clamped
    bilinear samples) against the oracle's FP64 transcription of the same algorithm."""
    _say(f"[sat method] case={name}")
    cfg, g = CASES[name]()
    plan = gpu_plan_factory(g)
    plan.set_method("sat")
    vol = volume(g)
    err = rel_l2(plan.forward(vol), oracle.fp(g, vol, "sat"))
    y = sinogram(g)
    err_b = rel_l2(plan.back(y, kind="ref"), oracle.bp(g, y, "sat"))
    _say(f"  sat fp rel-L2 = {err:.3e}, sat bp rel-L2 = {err_b:.3e}")
    assert err <= FP_TOL and err_b <= BP_TOL


def test_sat_method_clinical_view(oracle, gpu_plan_factory):
    """Clinical sizes are where whole-plane FP32 integral images fail (entries ~1e5);
    the compensated tables keep the four-corner differences accurate."""
    cfg, g = clinical()
    rng = np.random.default_rng(9)
    vol = rng.uniform(0, 2, g.vol_shape).astype(np.float32)
    plan = gpu_plan_factory(g)
    plan.set_method("sat")
    got = plan.forward(vol, [0])
    ref = oracle.fp(g.subset([0]), vol, "direct")
    err = rel_l2(got, ref)
    _say(f"  clinical sat fp view 0 rel-L2 = {err:.3e}")
    assert err <= FP_TOL


def test_sat_method_os_sart(oracle, gpu_plan_factory):
    from ctrecon_amd.plan import DeviceArray

    cfg, g = small()
    plan = gpu_plan_factory(g)
    plan.set_method("sat")
    vol = volume(g)
    b = plan.forward(vol)
    dx = DeviceArray(g.vol_shape)
    dx.zero()
    res = plan.os_sart(dx, DeviceArray.from_host(b), 10, 2, "ref")
    xr, res_r = oracle.os_sart(g, b, osNum=10, iterNum=2, bp_kind="ref")
    err = rel_l2(dx.get(), xr)
    _say(f"  sat os-sart iterate rel-L2 = {err:.3e}")
    assert err <= 1e-4

"""GPU parity of the BF16-storage projectors and solvers.

The BF16 kernels read BF16 inputs, accumulate in FP32 and (forward projector) round
the ray sums to BF16.  Bars:
  * conversions: bit-exact against numpy round-to-nearest-even;
  * forward: every ray within one BF16 ulp of the FP64 oracle's projection of the
    same BF16-rounded volume (the FP32 sum may move a value across a rounding
    boundary, never further);
  * back: FP32 output within 1e-5 relative L2 of the FP64 oracle's back projection
    of the same BF16-rounded sinogram (the FP32 bar);
  * OS-SART / CG with BF16 storage: the error against the FP64 oracle's iterate is
    reported and bounded by the storage precision (rel L2 <= 1e-2), and the FP32
    master copy keeps the data residual falling to the FP32 path's level.
"""

import numpy as np
import pytest

from bf16 import quantize, round_bits, ulp, widen
from cases import CASES, small, sinogram, volume
from conftest import rel_l2

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def test_conversions_bit_exact(gpu_plan_factory):
    from ctrecon_amd.plan import DeviceArray, from_bf16, to_bf16

    _say("[bf16 conversions]")
    gpu_plan_factory  # noqa: B018  (fails the test without a GPU)
    rng = np.random.default_rng(1)
    a = (rng.standard_normal(1 << 16) * rng.choice([1e-3, 1.0, 1e3], 1 << 16)).astype(np.float32)
    # exact ties and values around them
    t = (1 + (2 * np.arange(512) + 1) * 2.0**-8).astype(np.float32)
    a = np.concatenate([a, t, -t, np.nextafter(t, 2), np.nextafter(t, 0)]).astype(np.float32)
    d = DeviceArray.from_host(a)
    h = to_bf16(d).get()
    assert np.array_equal(h, round_bits(a))
    back = from_bf16(DeviceArray.from_host(h)).get()
    assert np.array_equal(back, widen(h))


@pytest.mark.parametrize("method", ["direct", "sat"])
@pytest.mark.parametrize("name", list(CASES))
def test_fp_bf16_parity(name, method, oracle, gpu_plan_factory):
    _say(f"[fp bf16] case={name} method={method}")
    cfg, g = CASES[name]()
    plan = gpu_plan_factory(g)
    plan.set_method(method)
    vol = volume(g)
    q = quantize(vol)
    got = plan.forward_bf16(vol)
    ref = oracle.fp(g, q, "sat")
    d = np.abs(got.astype(np.float64) - ref)
    # absolute slack for rays whose exact sum is ~0, where the FP32 accumulation
    # error (relative to the summands, not the sum) exceeds the BF16 spacing
    tiny = 1e-6 * float(np.max(np.abs(ref)))
    big = np.abs(ref) > tiny
    worst = float(np.max(d[big] / ulp(ref[big])))
    e_q = rel_l2(got, ref)
    e_f = rel_l2(got, oracle.fp(g, vol, "sat"))
    _say(f"  max |gpu - oracle(bf16 vol)| = {worst:.3f} BF16 ulp (rays above {tiny:.1e}); "
         f"rel-L2 vs oracle on the BF16 volume {e_q:.3e}, on the FP32 volume {e_f:.3e}")
    assert np.all(d <= ulp(ref) * 1.0000001 + tiny)
    assert e_q <= 4e-3


@pytest.mark.parametrize("method", ["direct", "sat"])
@pytest.mark.parametrize("name", list(CASES))
def test_bp_bf16_parity(name, method, oracle, gpu_plan_factory):
    _say(f"[bp-ref bf16] case={name} method={method}")
    cfg, g = CASES[name]()
    plan = gpu_plan_factory(g)
    plan.set_method(method)
    y = sinogram(g)
    got = plan.back_bf16(y, kind="ref")
    ref = oracle.bp(g, quantize(y), "sat")
    err = rel_l2(got, ref)
    _say(f"  rel-L2 vs oracle on the BF16 sinogram {err:.3e}; vs the FP32 sinogram "
         f"{rel_l2(got, oracle.bp(g, y, 'sat')):.3e}")
    assert err <= 1e-5


@pytest.mark.parametrize("name", list(CASES))
def test_bp_matched_bf16_parity(name, oracle, gpu_plan_factory):
    _say(f"[bp-matched bf16] case={name}")
    cfg, g = CASES[name]()
    plan = gpu_plan_factory(g)
    y = sinogram(g)
    got = plan.back_bf16(y, kind="matched")
    ref = oracle.bp(g, quantize(y), "adjoint")
    err = rel_l2(got, ref)
    _say(f"  rel-L2 vs oracle adjoint on the BF16 sinogram {err:.3e}")
    assert err <= 1e-5


@pytest.mark.parametrize("kind", ["ref", "matched"])
def test_os_sart_bf16(kind, oracle, gpu_plan_factory):
    from ctrecon_amd.plan import DeviceArray

    _say(f"[os-sart bf16] bp={kind}")
    cfg, g = small()
    plan = gpu_plan_factory(g)
    vol = volume(g)
    b = plan.forward(vol)
    xr, res_r = oracle.os_sart(g, b, osNum=10, iterNum=3, bp_kind=kind)
    out = {}
    for storage in ("f32", "bf16"):
        plan.set_storage(storage)
        dx = DeviceArray(g.vol_shape)
        dx.zero()
        res = plan.os_sart(dx, DeviceArray.from_host(b), 10, 3, kind)
        out[storage] = (rel_l2(dx.get(), xr), res)
    plan.set_storage("f32")
    _say(f"  iterate rel-L2 vs FP64 oracle: f32 {out['f32'][0]:.3e}, bf16 {out['bf16'][0]:.3e}; "
         f"residuals oracle {res_r} f32 {out['f32'][1]} bf16 {out['bf16'][1]}")
    assert out["f32"][0] <= 1e-4
    assert out["bf16"][0] <= 1e-2
    assert np.allclose(out["bf16"][1], res_r, rtol=2e-2)


@pytest.mark.parametrize("kind", ["ref", "matched"])
def test_conjgrad_bf16(kind, oracle, gpu_plan_factory):
    from ctrecon_amd.plan import DeviceArray

    _say(f"[cg bf16] bp={kind}")
    cfg, g = small()
    plan = gpu_plan_factory(g)
    vol = volume(g)
    b = plan.forward(vol)
    xr, res_r = oracle.conjgrad(g, b, iterNum=3, bp_kind=kind)
    plan.set_storage("bf16")
    dx = DeviceArray(g.vol_shape)
    dx.zero()
    res = plan.conjgrad(dx, DeviceArray.from_host(b), 3, kind)
    plan.set_storage("f32")
    err = rel_l2(dx.get(), xr)
    _say(f"  iterate rel-L2 vs FP64 oracle {err:.3e}; |r| oracle {res_r} bf16 {res}")
    assert err <= 1e-2


def test_bf16_master_copy_keeps_converging(gpu_plan_factory):
    """With BF16 storage the update is applied to the FP32 master iterate, so small
    late updates are not rounded away: the data residual follows the FP32 path."""
    from ctrecon_amd.plan import DeviceArray

    _say("[bf16 master copy] small, 10 subsets, 30 sweeps")
    cfg, g = small()
    plan = gpu_plan_factory(g)
    b = plan.forward(volume(g))
    res = {}
    for storage in ("f32", "bf16"):
        plan.set_storage(storage)
        dx = DeviceArray(g.vol_shape)
        dx.zero()
        res[storage] = plan.os_sart(dx, DeviceArray.from_host(b), 10, 30, "ref")
    plan.set_storage("f32")
    _say(f"  residual after 1/10/30 sweeps: f32 {res['f32'][[0, 9, 29]]} "
         f"bf16 {res['bf16'][[0, 9, 29]]}")
    assert res["bf16"][29] < res["bf16"][9]
    assert res["bf16"][29] <= 1.05 * res["f32"][29] + 1e-4

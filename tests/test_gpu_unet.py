"""GPU parity of the warm-start network's BF16-MFMA forward and backward kernels
against the FP64 oracle (oracle/unet_oracle.py) with BF16 rounding emulated where the
GPU stores BF16.  What remains is FP32-vs-FP64 accumulation and the BF16 rounding
boundaries it moves, so the bars are: forward output within 1e-2 relative L2, loss
within 1e-3, each layer's weight gradient within 3e-2 relative L2."""

import numpy as np
import pytest

from bf16 import quantize
from conftest import rel_l2

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def _q(a):
    return quantize(np.asarray(a, np.float32)).astype(np.float64)


def _inputs(shape, seed):
    rng = np.random.default_rng(seed)
    x = rng.uniform(0, 2, shape).astype(np.float32)
    d = (rng.normal(0, 0.05, shape)).astype(np.float32)
    t = (rng.normal(0, 0.05, shape)).astype(np.float32)
    return x, d, t


def _params(seed):
    from ctrecon_amd.unet import init_params

    p = init_params(seed)
    p += np.random.default_rng(seed + 100).normal(0, 0.01, p.size).astype(np.float32)
    return p


SHAPES = [(8, 8, 32), (12, 8, 64)]
SCALES = (1.0, 20.0, 0.05)


@pytest.mark.parametrize("shape", SHAPES)
def test_forward_parity(shape, gpu_plan_factory):
    from ctrecon_amd.plan import DeviceArray
    from ctrecon_amd.unet import UNet
    from oracle import unet_oracle as U

    _say(f"[unet forward] grid={shape}")
    x, d, _ = _inputs(shape, 0)
    p = _params(1)
    net = UNet(p, SCALES)
    out = DeviceArray(shape)
    net.forward(DeviceArray.from_host(x), DeviceArray.from_host(d), out)
    got = out.get()
    ref, _ = U.forward(p, x.astype(np.float64), d.astype(np.float64), SCALES, q=_q)
    err = rel_l2(got, ref)
    _say(f"  rel-L2 vs FP64 oracle (BF16 storage emulated) = {err:.3e}; "
         f"vs unrounded FP64 {rel_l2(got, U.forward(p, x, d, SCALES)[0]):.3e}")
    assert err <= 1e-2


@pytest.mark.parametrize("shape", SHAPES)
def test_gradient_parity(shape, gpu_plan_factory):
    from ctrecon_amd.plan import DeviceArray
    from ctrecon_amd.unet import LAYERS, UNet
    from oracle import unet_oracle as U

    _say(f"[unet gradient] grid={shape}")
    x, d, t = _inputs(shape, 2)
    p = _params(3)
    net = UNet(p, SCALES)
    loss, g = net.loss_grad(DeviceArray.from_host(x), DeviceArray.from_host(d),
                            DeviceArray.from_host(t))
    loss_r, g_r = U.loss_grad(p, x.astype(np.float64), d.astype(np.float64),
                              t.astype(np.float64), SCALES, q=_q)
    _say(f"  loss gpu {loss:.8e} oracle {loss_r:.8e}")
    assert abs(loss - loss_r) <= 1e-3 * abs(loss_r)
    o = 0
    worst = 0.0
    for l, (ci, co) in enumerate(LAYERS):
        nw = 27 * ci * co
        ew = rel_l2(g[o:o + nw], g_r[o:o + nw])
        eb = rel_l2(g[o + nw:o + nw + co], g_r[o + nw:o + nw + co])
        _say(f"  layer {l} ({ci}->{co}): dW rel-L2 {ew:.2e}, db rel-L2 {eb:.2e}")
        worst = max(worst, ew, eb)
        o += nw + co
    eh = rel_l2(g[o:], g_r[o:])
    _say(f"  head: rel-L2 {eh:.2e}")
    assert worst <= 3e-2 and eh <= 3e-2


def test_training_reduces_loss(gpu_plan_factory):
    from ctrecon_amd.plan import DeviceArray
    from ctrecon_amd.unet import UNet

    _say("[unet training] 120 Adam steps on one sample")
    shape = (16, 16, 32)
    rng = np.random.default_rng(5)
    x = rng.uniform(0, 2, shape).astype(np.float32)
    d = rng.normal(0, 0.05, shape).astype(np.float32)
    t = (0.5 * d + 0.02 * (x - 1.0)).astype(np.float32)  # a learnable map of the inputs
    net = UNet(None, (1.0, 20.0, 0.05))
    dx, dd, dt = (DeviceArray.from_host(a) for a in (x, d, t))
    losses = [net.train_step(dx, dd, dt, 2e-3) for _ in range(120)]
    _say(f"  loss {losses[0]:.4e} -> {losses[59]:.4e} (60 steps) -> {losses[-1]:.4e}")
    assert np.all(np.isfinite(losses))
    assert losses[-1] < 0.3 * losses[0]
    assert losses[-1] < losses[59]  # still falling

"""The learned warm-start network over the C ABI (ctr_unet_*).

:class:`UNet` owns a network on one GPU: FP32 master parameters, BF16 copies used by
the MFMA convolution kernels, and Adam state.  Parameter files are ``.npz`` archives
holding ``params`` (float32, :data:`LAYERS` order) and ``scales`` (s_x, s_d, s_out).
"""

from __future__ import annotations

import ctypes

import numpy as np

from ._lib import check, lib
from .plan import DeviceArray

__all__ = ["LAYERS", "UNet", "init_params", "num_params", "save", "load"]

# (Cin, Cout) of the ten 3x3x3 convolutions, in parameter order:
# e1a e1b | e2a e2b | e3a e3b | d2a d2b | d1a d1b, then the 16 -> 1 head.
LAYERS = [(8, 16), (16, 16), (16, 32), (32, 32), (32, 64), (64, 64), (96, 32), (32, 32),
          (48, 16), (16, 16)]


def num_params() -> int:
    return int(sum(27 * ci * co + co for ci, co in LAYERS) + 17)


def init_params(seed: int = 0) -> np.ndarray:
    """He-normal convolution weights, zero biases, small head weights."""
    rng = np.random.default_rng(seed)
    out = []
    for ci, co in LAYERS:
        fan_in = 27 * (2 if ci == 8 else ci)  # the first layer has 2 live input channels
        out.append(rng.normal(0.0, np.sqrt(2.0 / fan_in), 27 * co * ci))
        out.append(np.zeros(co))
    out.append(rng.normal(0.0, 0.1 / 4.0, 16))
    out.append(np.zeros(1))
    p = np.concatenate(out).astype(np.float32)
    assert p.size == num_params()
    return p


def save(path, params, scales):
    np.savez(path, params=np.asarray(params, np.float32), scales=np.asarray(scales, np.float32))


def load(path):
    with np.load(path, allow_pickle=False) as z:
        return z["params"].astype(np.float32), tuple(float(s) for s in z["scales"])


class UNet:
    """A network instance on one device."""

    def __init__(self, params=None, scales=(1.0, 1.0, 1.0), device=0):
        self.device = device
        h = ctypes.c_void_p()
        check(lib().ctr_unet_create(device, ctypes.byref(h)))
        self.handle = h.value
        self.set_params(init_params() if params is None else params)
        self.set_scales(*scales)

    @classmethod
    def from_file(cls, path, device=0):
        p, s = load(path)
        return cls(p, s, device)

    def save(self, path):
        save(path, self.get_params(), self.scales)

    def set_params(self, params):
        p = np.ascontiguousarray(params, dtype=np.float32)
        assert p.size == num_params(), (p.size, num_params())
        check(lib().ctr_unet_set_params(self.handle, p.ctypes.data, p.size))

    def get_params(self) -> np.ndarray:
        p = np.zeros(num_params(), np.float32)
        check(lib().ctr_unet_get_params(self.handle, p.ctypes.data, p.size))
        return p

    def set_scales(self, s_x, s_d, s_out):
        check(lib().ctr_unet_set_scales(self.handle, float(s_x), float(s_d), float(s_out)))
        self.scales = (float(s_x), float(s_d), float(s_out))

    def forward(self, d_x: DeviceArray, d_delta: DeviceArray, d_out: DeviceArray, stream=None):
        Y, X, Z = d_x.shape
        check(lib().ctr_unet_forward(self.handle, d_x.ptr, d_delta.ptr, d_out.ptr, Y, X, Z,
                                     stream))

    def loss_grad(self, d_x, d_delta, d_target):
        Y, X, Z = d_x.shape
        g = np.zeros(num_params(), np.float32)
        loss = ctypes.c_double()
        check(lib().ctr_unet_loss_grad(self.handle, d_x.ptr, d_delta.ptr, d_target.ptr, Y, X, Z,
                                       g.ctypes.data, ctypes.byref(loss)))
        return loss.value, g

    def train_step(self, d_x, d_delta, d_target, lr) -> float:
        Y, X, Z = d_x.shape
        loss = ctypes.c_double()
        check(lib().ctr_unet_train_step(self.handle, d_x.ptr, d_delta.ptr, d_target.ptr, Y, X, Z,
                                        float(lr), ctypes.byref(loss)))
        return loss.value

    def close(self):
        if getattr(self, "handle", None):
            lib().ctr_unet_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

"""Pins of the FP64 network oracle: the convolution against its definition, the
pooling / upsampling adjoints, and the hand-written gradient against central finite
differences."""

import numpy as np

from oracle import unet_oracle as U


def test_conv3_matches_definition():
    rng = np.random.default_rng(0)
    x = rng.standard_normal((3, 4, 5, 2))
    W = rng.standard_normal((27, 3, 2))
    b = rng.standard_normal(3)
    got = U.conv3(x, W, b)
    want = np.zeros((3, 4, 5, 3))
    for y in range(3):
        for xx in range(4):
            for z in range(5):
                for t in range(27):
                    dy, dx, dz = t // 9 - 1, (t // 3) % 3 - 1, t % 3 - 1
                    yy, xs, zz = y + dy, xx + dx, z + dz
                    if 0 <= yy < 3 and 0 <= xs < 4 and 0 <= zz < 5:
                        want[y, xx, z] += W[t] @ x[yy, xs, zz]
                want[y, xx, z] += b
    assert np.allclose(got, want, rtol=0, atol=1e-12)


def test_pool_and_upsample_adjoints():
    rng = np.random.default_rng(1)
    x = rng.standard_normal((4, 6, 8, 3))
    y = rng.standard_normal((2, 3, 4, 3))
    assert np.isclose(np.sum(U.pool(x) * y), np.sum(x * U.pool_adj(y)), rtol=1e-13)
    assert np.isclose(np.sum(U.up(y) * x), np.sum(y * U.up_adj(x)), rtol=1e-13)
    # pooling averages: a constant field is unchanged
    assert np.allclose(U.pool(np.full((4, 4, 4, 2), 3.0)), 3.0)


def test_gradient_matches_finite_differences():
    from ctrecon_amd.unet import init_params

    rng = np.random.default_rng(2)
    p = init_params(3).astype(np.float64)
    p += rng.normal(0, 0.02, p.size)  # non-zero biases
    shape = (4, 4, 8)
    x = rng.uniform(0, 2, shape)
    d = rng.normal(0, 1, shape)
    t = rng.normal(0, 1, shape)
    scales = (1.0, 0.5, 2.0)
    loss, g = U.loss_grad(p, x, d, t, scales)
    # parameter indices in every layer: weights and biases separately, and the head
    idx, o = [], 0
    for ci, co in U.LAYERS:
        nw = 27 * co * ci
        idx += list(rng.integers(o, o + nw, 3)) + list(rng.integers(o + nw, o + nw + co, 2))
        o += nw + co
    idx += list(range(o, o + 17, 4)) + [o + 16]
    idx = np.array(idx)
    eps = 1e-6
    ok = 0
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += eps
        pm[i] -= eps
        fd = (U.loss_grad(pp, x, d, t, scales)[0] - U.loss_grad(pm, x, d, t, scales)[0]) / (2 * eps)
        if abs(fd - g[i]) <= 1e-5 * max(abs(fd), 1e-3 * abs(loss)):
            ok += 1
    # a ReLU kink inside [p - eps, p + eps] can spoil a sample; a wrong adjoint spoils most
    assert ok >= 0.9 * len(idx), (ok, len(idx))

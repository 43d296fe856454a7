"""FP64 numpy restatement of the warm-start network and its gradient (TEST
INFRASTRUCTURE ONLY: imported by tests/, never by the product).

The network has no counterpart in the reference; this file restates it from its
definition (the architecture in ctrecon_amd/csrc/unet.cpp): 3x3x3 "same" convolutions
with zero padding, ReLU, 2x2x2 average pooling, nearest 2x upsampling, channel
concatenation of the skips, a 1x1x1 head, and loss sum((h - t / s_out)^2) / (2 V).
The backward pass is written out by hand (the adjoint of every step).

Pinned by: the convolution against a brute-force loop over its definition, pooling /
upsampling adjointness, and the whole gradient against central finite differences
(tests/test_unet_oracle.py).

q: an optional rounding function applied where the GPU stores BF16 (weights, the packed
input, every activation and activation gradient), so that the GPU path can be compared
at the precision it stores; identity by default.
"""

from __future__ import annotations

import numpy as np

LAYERS = [(8, 16), (16, 16), (16, 32), (32, 32), (32, 64), (64, 64), (96, 32), (32, 32),
          (48, 16), (16, 16)]


def _ident(a):
    return a


def unpack(params):
    p = np.asarray(params, np.float64)
    o = 0
    Ws, bs = [], []
    for ci, co in LAYERS:
        n = 27 * co * ci
        Ws.append(p[o:o + n].reshape(27, co, ci))
        o += n
        bs.append(p[o:o + co])
        o += co
    hw, hb = p[o:o + 16], p[o + 16]
    assert o + 17 == p.size
    return Ws, bs, hw, hb


def pack(Ws, bs, hw, hb):
    out = []
    for W, b in zip(Ws, bs):
        out += [W.ravel(), b.ravel()]
    out += [np.ravel(hw), np.atleast_1d(hb)]
    return np.concatenate(out)


def _taps():
    for t in range(27):
        yield t, t // 9 - 1, (t // 3) % 3 - 1, t % 3 - 1


def conv3(x, W, b):
    """x [Y, X, Z, Cin], W [27, Cout, Cin] -> [Y, X, Z, Cout] (+ b)."""
    Y, X, Z, _ = x.shape
    xp = np.pad(x, ((1, 1), (1, 1), (1, 1), (0, 0)))
    out = np.zeros((Y, X, Z, W.shape[1]))
    for t, dy, dx, dz in _taps():
        out += xp[1 + dy:1 + dy + Y, 1 + dx:1 + dx + X, 1 + dz:1 + dz + Z] @ W[t].T
    return out + (0.0 if b is None else b)


def conv3_bwd(x, W, dP):
    """Gradients of conv3 for the pre-activation gradient dP: (dX, dW, db)."""
    Y, X, Z, Cin = x.shape
    xp = np.pad(x, ((1, 1), (1, 1), (1, 1), (0, 0)))
    dxp = np.zeros_like(xp)
    dW = np.zeros_like(W)
    dPf = dP.reshape(-1, dP.shape[-1])
    for t, dy, dx, dz in _taps():
        sl = (slice(1 + dy, 1 + dy + Y), slice(1 + dx, 1 + dx + X), slice(1 + dz, 1 + dz + Z))
        dW[t] = dPf.T @ xp[sl].reshape(-1, Cin)
        dxp[sl] += dP @ W[t]
    return dxp[1:-1, 1:-1, 1:-1], dW, dPf.sum(axis=0)


def pool(x):
    Y, X, Z, C = x.shape
    return x.reshape(Y // 2, 2, X // 2, 2, Z // 2, 2, C).mean(axis=(1, 3, 5))


def up(x):
    return x.repeat(2, axis=0).repeat(2, axis=1).repeat(2, axis=2)


def pool_adj(d):
    return up(d) / 8.0


def up_adj(d):
    return pool(d) * 8.0


def forward(params, x, delta, scales=(1.0, 1.0, 1.0), q=_ident, head_scale=None):
    """Activations of the network on volumes x, delta [Y, X, Z]; returns (h, acts) with
    h the head output times head_scale (default s_out)."""
    Ws, bs, hw, hb = unpack(params)
    Wq = [q(W) for W in Ws]
    s_x, s_d, s_out = scales
    relu = lambda a: np.maximum(a, 0.0)  # noqa: E731
    A = {}
    inp = np.zeros(x.shape + (8,))
    inp[..., 0] = x * s_x
    inp[..., 1] = delta * s_d
    A["in0"] = q(inp)
    A["a0"] = q(relu(conv3(A["in0"], Wq[0], bs[0])))
    A["s1"] = q(relu(conv3(A["a0"], Wq[1], bs[1])))
    A["p1"] = q(pool(A["s1"]))
    A["a2"] = q(relu(conv3(A["p1"], Wq[2], bs[2])))
    A["s2"] = q(relu(conv3(A["a2"], Wq[3], bs[3])))
    A["p2"] = q(pool(A["s2"]))
    A["a4"] = q(relu(conv3(A["p2"], Wq[4], bs[4])))
    A["b3"] = q(relu(conv3(A["a4"], Wq[5], bs[5])))
    A["cat2"] = np.concatenate([A["s2"], up(A["b3"])], axis=-1)
    A["a6"] = q(relu(conv3(A["cat2"], Wq[6], bs[6])))
    A["d2"] = q(relu(conv3(A["a6"], Wq[7], bs[7])))
    A["cat1"] = np.concatenate([A["s1"], up(A["d2"])], axis=-1)
    A["a8"] = q(relu(conv3(A["cat1"], Wq[8], bs[8])))
    A["d1"] = q(relu(conv3(A["a8"], Wq[9], bs[9])))
    hs = s_out if head_scale is None else head_scale
    h = (A["d1"] @ hw + hb) * hs
    return h, A


def loss_grad(params, x, delta, target, scales=(1.0, 1.0, 1.0), q=_ident):
    """Loss sum((h - target / s_out)^2) / (2 V) of the unscaled head output and its
    gradient with respect to params (FP64)."""
    Ws, bs, hw, hb = unpack(params)
    Wq = [q(W) for W in Ws]
    s_out = scales[2]
    h, A = forward(params, x, delta, scales, q, head_scale=1.0)
    V = h.size
    e = h - target / s_out
    loss = float(np.sum(e * e) / (2.0 * V))
    dh = e / V
    gW = [None] * 10
    gb = [None] * 10
    g_hw = np.tensordot(A["d1"], dh, axes=([0, 1, 2], [0, 1, 2]))
    g_hb = float(dh.sum())
    mask = lambda y, d: np.where(y > 0, d, 0.0)  # noqa: E731

    def layer(l, inp, dP):
        dX, gW[l], gb[l] = conv3_bwd(inp, Wq[l], dP)
        return q(dX)

    dP = q(mask(A["d1"], dh[..., None] * hw))
    dP = q(mask(A["a8"], layer(9, A["a8"], dP)))
    dcat1 = layer(8, A["cat1"], dP)
    dP = q(mask(A["d2"], q(up_adj(dcat1[..., 16:]))))
    dP = q(mask(A["a6"], layer(7, A["a6"], dP)))
    dcat2 = layer(6, A["cat2"], dP)
    dP = q(mask(A["b3"], q(up_adj(dcat2[..., 32:]))))
    dP = q(mask(A["a4"], layer(5, A["a4"], dP)))
    dp2 = layer(4, A["p2"], dP)
    dP = q(mask(A["s2"], q(dcat2[..., :32] + pool_adj(dp2))))
    dP = q(mask(A["a2"], layer(3, A["a2"], dP)))
    dp1 = layer(2, A["p1"], dP)
    dP = q(mask(A["s1"], q(dcat1[..., :16] + pool_adj(dp1))))
    dP = q(mask(A["a0"], layer(1, A["a0"], dP)))
    layer(0, A["in0"], dP)
    return loss, pack(gW, gb, g_hw, g_hb)

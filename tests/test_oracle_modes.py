"""CPU tests that pin the oracle's restatements of the reference's other projector
modes (no GPU): the pseudo distance-driven pair (prjMode 1 / 3 forward, 1 / 2 back)
and the FP64 branchless forward projector (prjMode 2).

Pins:
- pseudo-DD forward: the path length of a ray through a slab of ones (closed form),
  an affine volume (bilinear interpolation reproduces it exactly, so the ray sum is
  the value at the slab's mid plane times the path length), and a random masked,
  off-centre volume against scipy's border-addressed linear interpolation
  (``map_coordinates(order=1, mode='grid-constant')``) at sample points found from the
  source -> cell line;
- pseudo-DD back: a random sinogram against the same scipy routine at detector
  coordinates found with atan2 fan angles and similar triangles (the oracle uses the
  reference's asin / calDetParas form), and the mask semantics (only bytes equal to 1);
- FP64 branchless forward: equal to the mode-0 footprint model with the image centre
  moved by half a cell along the in-plane axis across the rays and along z (views of
  one marching direction at a time), the direct mode-0 restatement being pinned
  elsewhere.
"""

import copy
import math

import numpy as np
import pytest
from scipy import ndimage

from conftest import rel_l2
from cases import ragged, sinogram, tiny

from ctrecon_amd import geometry as G


def _view(g, p):
    th = float(np.float64(g.angles[p])) % (2 * math.pi)
    c, s = math.cos(th), math.sin(th)
    S = np.array([g.x0 * c - g.y0 * s, g.x0 * s + g.y0 * c, g.z0 + float(g.zshifts[p])])
    xds, yds = g.xds.astype(np.float64), g.yds.astype(np.float64)
    Dx, Dy = xds * c - yds * s, xds * s + yds * c
    Dz = g.zds.astype(np.float64) + float(g.zshifts[p])
    return c, s, S, Dx, Dy, Dz


def _oc(g):
    """objCntIdx from the centre as the C ABI receives it (float32)."""
    cx, cy, cz = (float(np.float32(c)) for c in (g.imgXCenter, g.imgYCenter, g.imgZCenter))
    return ((g.XN - 1) / 2 - cx / g.dx, (g.YN - 1) / 2 - cy / g.dx, (g.ZN - 1) / 2 - cz / g.dz)


def _pdd_fp_scipy(g, vol):
    """Ray sums of border-addressed bilinear samples on the slice centre planes."""
    ocx, ocy, ocz = _oc(g)
    mv = vol.astype(np.float64) * g.mask.reshape(g.YN, g.XN)[:, :, None]
    out = np.zeros(g.prj_shape)
    for p in range(g.PN):
        c, s, S, Dx, Dy, Dz = _view(g, p)
        ex, ey = Dx - S[0], Dy - S[1]
        ez = Dz - S[2]
        L = np.sqrt(ex[:, None] ** 2 + ey[:, None] ** 2 + ez[None, :] ** 2)
        acc = np.zeros((g.DNU, g.DNV))
        if abs(c) <= abs(s):
            for ii in range(g.XN):
                t = ((ii - ocx) * g.dx - S[0]) / ex
                yi = (S[1] + t * ey) / g.dx + ocy
                zi = (S[2] + t[:, None] * ez[None, :]) / g.dz + ocz
                pts = np.stack([np.broadcast_to(yi[:, None], zi.shape).ravel(), zi.ravel()])
                acc += ndimage.map_coordinates(mv[:, ii, :], pts, order=1,
                                               mode="grid-constant").reshape(acc.shape)
            out[p] = acc * g.dx * L / np.abs(ex)[:, None]
        else:
            for jj in range(g.YN):
                t = ((jj - ocy) * g.dx - S[1]) / ey
                xi = (S[0] + t * ex) / g.dx + ocx
                zi = (S[2] + t[:, None] * ez[None, :]) / g.dz + ocz
                pts = np.stack([np.broadcast_to(xi[:, None], zi.shape).ravel(), zi.ravel()])
                acc += ndimage.map_coordinates(mv[jj, :, :], pts, order=1,
                                               mode="grid-constant").reshape(acc.shape)
            out[p] = acc * g.dx * L / np.abs(ey)[:, None]
    return out


def _bounds(c):
    c = np.asarray(c, np.float64)
    b = np.empty(c.size + 1)
    b[1:-1] = 0.5 * (c[:-1] + c[1:])
    b[0] = 1.5 * c[0] - 0.5 * c[1]
    b[-1] = 1.5 * c[-1] - 0.5 * c[-2]
    return b


def _pdd_bp_scipy(g, prj):
    """Sum over views of dx |d| / |d_x or d_y| times the border-addressed bilinear sample
    of the sinogram where the source -> voxel ray meets the detector."""
    ocx, ocy, ocz = _oc(g)
    x0, y0, z0 = g.x0, g.y0, g.z0
    bx, by, bz = _bounds(g.xds), _bounds(g.yds), _bounds(g.zds)

    def ang(ax, ay, bx_, by_):  # signed angle from a to b
        return np.arctan2(ax * by_ - ay * bx_, ax * bx_ + ay * by_)

    beta0 = ang(-x0, -y0, bx[0] - x0, by[0] - y0)
    dbeta = ang(bx[0] - x0, by[0] - y0, bx[-1] - x0, by[-1] - y0) / g.DNU
    ddv = (bz[-1] - bz[0]) / g.DNV
    S2D = math.hypot(float(g.xds[0]) - x0, float(g.yds[0]) - y0)
    Y, X, Z = np.meshgrid(np.arange(g.YN), np.arange(g.XN), np.arange(g.ZN), indexing="ij")
    Px, Py, Pz = (X - ocx) * g.dx, (Y - ocy) * g.dx, (Z - ocz) * g.dz
    out = np.zeros(g.vol_shape)
    for p in range(g.PN):
        c, s, S, *_ = _view(g, p)
        dX, dY, dZ = Px - S[0], Py - S[1], Pz - S[2]
        rxy = np.hypot(dX, dY)
        alpha = ang(-S[0], -S[1], dX, dY)
        iu = (alpha - beta0) / dbeta - 0.5
        iz = (z0 + dZ * S2D / rxy - bz[0]) / ddv - 0.5
        w = dX if abs(S[0]) >= abs(S[1]) else dY
        cosv = g.dx * np.sqrt(rxy ** 2 + dZ ** 2) / np.abs(w)
        val = ndimage.map_coordinates(prj[p].astype(np.float64), np.stack([iu.ravel(), iz.ravel()]),
                                      order=1, mode="grid-constant").reshape(iu.shape)
        out += val * cosv
    m = g.mask.reshape(g.YN, g.XN)
    out[m != 1] = 0.0
    return out


def _slab(xmarch):
    """Thin slab across the rays of the chosen marching direction, wide along the other
    axes, narrow fan and cone, so every sample of every ray is interior."""
    cfg = G.helical_config(sid=400, sdd=700, DNU=9, DNV=4, col_width=1.0, col_height=1.0,
                           views_per_turn=48, turns=1, feed_mm_per_turn=2.0, N_xy=64, N_z=48,
                           voxel_xy=1.0, voxel_z=1.0)
    g = G.dd3_geometry(cfg)
    views = [p for p in range(g.PN)
             if (abs(math.cos(float(g.angles[p]))) <= abs(math.sin(float(g.angles[p])))) == xmarch
             and abs(abs(math.cos(float(g.angles[p]))) - abs(math.sin(float(g.angles[p])))) > 0.3]
    g = g.subset(views)
    if xmarch:
        g.XN = 6
    else:
        g.YN = 6
    g.mask = np.ones(g.XN * g.YN, np.uint8)
    return g


@pytest.mark.parametrize("xmarch", [True, False])
def test_pdd_fp_slab_path_length(xmarch, oracle):
    """Ones in a slab of N slices: every ray sum is N dx |e| / |e_along| (closed form)."""
    g = _slab(xmarch)
    got = oracle.fp(g, np.ones(g.vol_shape, np.float32), "pdd")
    N = g.XN if xmarch else g.YN
    for p in range(g.PN):
        c, s, S, Dx, Dy, Dz = _view(g, p)
        e = np.stack(np.broadcast_arrays((Dx - S[0])[:, None], (Dy - S[1])[:, None],
                                         (Dz - S[2])[None, :]))
        L = np.sqrt((e ** 2).sum(0))
        ref = N * g.dx * L / np.abs(e[0] if xmarch else e[1])
        assert np.allclose(got[p], ref, rtol=1e-12)


@pytest.mark.parametrize("xmarch", [True, False])
def test_pdd_fp_affine_volume(xmarch, oracle):
    """Affine in the cell indices: the ray sum is the value where the ray crosses the
    slab's middle plane times the path length."""
    g = _slab(xmarch)
    ocx, ocy, ocz = _oc(g)
    a, b, c0 = 0.37, -0.21, 5.0
    Y, X, Z = np.meshgrid(np.arange(g.YN), np.arange(g.XN), np.arange(g.ZN), indexing="ij")
    lat = X if not xmarch else Y
    vol = (c0 + a * lat + b * Z).astype(np.float64)
    got = oracle.fp(g, vol.astype(np.float32), "pdd")
    vol32 = vol.astype(np.float32).astype(np.float64)
    assert np.allclose(vol32, vol, rtol=1e-6)
    N = g.XN if xmarch else g.YN
    for p in range(g.PN):
        c, s, S, Dx, Dy, Dz = _view(g, p)
        ex, ey, ez = Dx - S[0], Dy - S[1], Dz - S[2]
        if xmarch:
            t = (((N - 1) / 2 - ocx) * g.dx - S[0]) / ex
            li = (S[1] + t * ey) / g.dx + ocy
            along = np.abs(ex)[:, None]
        else:
            t = (((N - 1) / 2 - ocy) * g.dx - S[1]) / ey
            li = (S[0] + t * ex) / g.dx + ocx
            along = np.abs(ey)[:, None]
        zi = (S[2] + t[:, None] * ez[None, :]) / g.dz + ocz
        L = np.sqrt(ex[:, None] ** 2 + ey[:, None] ** 2 + ez[None, :] ** 2)
        ref = (c0 + a * li[:, None] + b * zi) * N * g.dx * L / along
        assert np.allclose(got[p], ref, rtol=2e-6)


@pytest.mark.parametrize("case", [tiny, ragged])
def test_pdd_fp_vs_scipy(case, oracle):
    cfg, g = case()
    vol = np.random.default_rng(3).uniform(0, 1, g.vol_shape).astype(np.float32)
    got = oracle.fp(g, vol, "pdd")
    ref = _pdd_fp_scipy(g, vol)
    assert rel_l2(got, ref) < 1e-12


@pytest.mark.parametrize("case", [tiny, ragged])
def test_pdd_bp_vs_scipy(case, oracle):
    cfg, g = case()
    y = sinogram(g)
    got = oracle.bp(g, y, "pdd")
    ref = _pdd_bp_scipy(g, y)
    assert rel_l2(got, ref) < 1e-11
    m = g.mask.reshape(g.YN, g.XN)
    assert np.all(got[m != 1] == 0) and np.any(got[m == 1] != 0)


def test_pdd_mask_byte_scales_forward(oracle):
    """The forward projector multiplies the volume by the mask byte (DD3_GPU.cu:766-774)."""
    cfg, g = tiny()
    vol = np.random.default_rng(4).uniform(0, 1, g.vol_shape).astype(np.float32)
    g2 = copy.copy(g)
    g2.mask = np.full_like(g.mask, 2)
    assert np.allclose(oracle.fp(g2, vol, "pdd"), 2 * oracle.fp(g, vol, "pdd"), rtol=1e-14)


def _march_views(g, xmarch):
    return [p for p in range(g.PN)
            if (abs(math.cos(float(np.float64(g.angles[p]) % (2 * math.pi))))
                <= abs(math.sin(float(np.float64(g.angles[p]) % (2 * math.pi))))) == xmarch]


@pytest.mark.parametrize("case", [tiny, ragged])
@pytest.mark.parametrize("xmarch", [True, False])
def test_fp_dp_is_half_cell_shifted_mode0(case, xmarch, oracle):
    """prjMode 2 samples its integral images at real / d + objCntIdx + 1 without the
    half-texel offset of linear filtering: the footprint model of mode 0 evaluated with
    objCntIdx + 0.5 along the axis across the rays and along z."""
    cfg, g = case()
    g = g.subset(_march_views(g, xmarch))
    vol = np.random.default_rng(5).uniform(0, 1, g.vol_shape).astype(np.float32)
    got = oracle.fp(g, vol, "dp")
    h = copy.copy(g)
    if xmarch:
        h.imgYCenter = g.imgYCenter - 0.5 * g.dx
    else:
        h.imgXCenter = g.imgXCenter - 0.5 * g.dx
    h.imgZCenter = g.imgZCenter - 0.5 * g.dz
    ref = oracle.fp(h, vol, "direct")
    # the shifted centre is passed as float32 like every geometry argument: exact for the
    # tiny case (centre 0, dx 4, dz 3), within ~1e-7 mm for the off-centre ragged one
    assert rel_l2(got, ref) < (1e-10 if case is tiny else 1e-6)
    # and it is not the unshifted model
    assert rel_l2(got, oracle.fp(g, vol, "direct")) > 1e-3


# ---------------------------------------------------------------- z-line back (prjMode 3)
def _zline_numpy(g, prj):
    """Sum over views of cosVal w(j) G(tD_j) at the lower faces j of the slices, G the
    exact integral of the piecewise-constant view over the fan-angle range of the
    column's extent and the rows below tD; slice k gets S(k+1) - S(k) for the whole
    16-slice blocks, 0 above them."""
    ocx, ocy, ocz = _oc(g)
    x0, y0, z0 = g.x0, g.y0, g.z0
    bx, by, bz = _bounds(g.xds), _bounds(g.yds), _bounds(g.zds)

    def ang(ax, ay, bx_, by_):
        return math.atan2(ax * by_ - ay * bx_, ax * bx_ + ay * by_)

    beta0 = ang(-x0, -y0, bx[0] - x0, by[0] - y0)
    dbeta = ang(bx[0] - x0, by[0] - y0, bx[-1] - x0, by[-1] - y0) / g.DNU
    ddv = (bz[-1] - bz[0]) / g.DNV
    S2D = math.hypot(float(g.xds[0]) - x0, float(g.yds[0]) - y0)
    NZ = 16 * (g.ZN // 16)
    zf = (np.arange(NZ + 1) - ocz) * g.dz  # slice centres 0..NZ
    cum = np.concatenate([np.zeros(prj.shape[:2] + (1,)), np.cumsum(prj, axis=2)], axis=2)
    grid = np.arange(g.DNV + 1)
    m = g.mask.reshape(g.YN, g.XN)
    out = np.zeros(g.vol_shape)
    for y in range(g.YN):
        for x in range(g.XN):
            if m[y, x] != 1 or NZ == 0:
                continue
            Px, Py = (x - ocx) * g.dx, (y - ocy) * g.dx
            S = np.zeros(NZ + 1)
            for p in range(g.PN):
                c, s, Sv, *_ = _view(g, p)
                dX, dY = Px - Sv[0], Py - Sv[1]
                s2v = math.hypot(dX, dY)
                if abs(c) <= abs(s):
                    L, R, cosv = (dX, dY - g.dx / 2), (dX, dY + g.dx / 2), g.dx * s2v / dX
                else:
                    L, R, cosv = (dX + g.dx / 2, dY), (dX - g.dx / 2, dY), g.dx * s2v / dY
                tL = (ang(-Sv[0], -Sv[1], *L) - beta0) / dbeta - 0.5
                tR = (ang(-Sv[0], -Sv[1], *R) - beta0) / dbeta - 0.5
                lo = min(max(min(tL, tR), -0.5), g.DNU - 0.5)
                hi = min(max(max(tL, tR), -0.5), g.DNU - 0.5)
                sgn = 1.0 if tR >= tL else -1.0
                dirZ = zf - Sv[2]
                w = np.hypot(dirZ, s2v) / s2v
                tD = (z0 + (dirZ - g.dz / 2) * S2D / s2v - bz[0]) / ddv - 0.5
                G = np.zeros(NZ + 1)
                for u in range(g.DNU):
                    wu = min(hi, u + 0.5) - max(lo, u - 0.5)
                    if wu > 0:
                        G += wu * np.interp(tD + 0.5, grid, cum[p, u])
                S += sgn * cosv * w * G
            out[y, x, :NZ] = S[1:] - S[:-1]
    return out


def _tall_tiny():
    cfg = G.helical_config(sid=60, sdd=100, DNU=12, DNV=4, col_width=4.0, col_height=3.0,
                           views_per_turn=8, turns=2, feed_mm_per_turn=6.0, N_xy=8, N_z=32,
                           voxel_xy=4.0, voxel_z=1.5)
    return cfg, G.dd3_geometry(cfg)


@pytest.mark.parametrize("case", [_tall_tiny, ragged])
def test_zline_vs_numpy(case, oracle):
    """ragged has ZN = 37: two whole 16-slice blocks, the last 5 slices stay 0."""
    cfg, g = case()
    y = sinogram(g)
    got = oracle.bp(g, y, "zline")
    ref = _zline_numpy(g, y.astype(np.float64))
    assert rel_l2(got, ref) < 1e-10
    NZ = 16 * (g.ZN // 16)
    assert np.all(got[:, :, NZ:] == 0)
    m = g.mask.reshape(g.YN, g.XN)
    assert np.all(got[m != 1] == 0) and np.any(got[m == 1] != 0)


def test_zline_fewer_than_16_slices_is_zero(oracle):
    cfg, g = tiny()  # ZN = 8: no whole block
    assert np.all(oracle.bp(g, sinogram(g), "zline") == 0)

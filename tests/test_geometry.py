"""Geometry (dd3.m restatement), configs and phantoms -- CPU only."""

import math

import numpy as np

from ctrecon_amd import geometry as G
from ctrecon_amd import phantom as P


def test_arc_detector_closed_form():
    """dd3.m:150-152: alpha_i = (i - (col_center+1)) w / sdd, xds = sdd sin, yds = sid - sdd cos."""
    cfg = G.config_clinical()
    g = G.dd3_geometry(cfg)
    i = np.arange(1, 889, dtype=np.float64)
    alpha = (i - ((888 - 1) / 2 + 1)) * 1.0239 / 949
    assert np.allclose(g.xds, 949 * np.sin(alpha), rtol=0, atol=2e-4)
    assert np.allclose(g.yds, 541 - 949 * np.cos(alpha), rtol=0, atol=2e-4)
    # a centred detector is symmetric about its middle
    assert np.allclose(g.xds, -g.xds[::-1], atol=2e-4)


def test_rows_and_pitch():
    cfg = G.config_clinical()
    g = G.dd3_geometry(cfg)
    j = np.arange(1, 65)
    assert np.allclose(g.zds, (j - 32.5) * 1.096439, atol=1e-5)  # dd3.m:159
    # 0.9 pitch with 64 x 0.625 mm at iso: 36 mm of table feed per turn (dd3.m:178)
    feed = g.zshifts[984] - g.zshifts[0]
    assert abs(feed - 36.0) < 1e-3
    assert abs((g.zshifts[-1] + g.zshifts[0])) < 1e-3  # travel centred on z = 0
    assert g.PN == 3936 and g.DNU == 888 and g.DNV == 64
    assert abs(g.dx - 0.7) < 1e-6 and abs(g.dz - 0.7) < 1e-6


def test_gpu_axis_swap():
    """The GPU path receives angles + pi/2 (dd3.m:247) and swapped offsets (:249-250)."""
    cfg = G.config_small()
    g = G.dd3_geometry(cfg)
    k = np.arange(g.PN)
    want = (k * 2 * math.pi / 360) + math.pi / 2
    assert np.allclose(g.angles, want, atol=1e-5)
    assert g.imgXCenter == 0 and g.imgYCenter == 0


def test_view_subset_matches_os_sart_indexing():
    # os_sart.m:15: views = viewSer(1+osIdx:osNum:end), one-based
    v = G.view_subset(3936, 3, 20)
    assert v[0] == 3 and v[1] == 23 and v[-1] == 3923 and v.size == 197
    assert G.view_subset(3936, 19, 20).size == 196


def test_mask_byte_layout():
    """The GPU reads mask[y*XN + x] from the first XN*YN bytes of int8([mask, mask'])
    (dd3.m:185, DD3_GPU.cu:1006): that is MATLAB's column-major mask(x, y)."""
    cfg = G.config_small()
    m = np.zeros((64, 64), np.int8)
    m[5, 9] = 1  # x = 5, y = 9
    g = G.dd3_geometry(cfg, mask=m)
    assert g.mask[9 * 64 + 5] == 1 and g.mask.sum() == 1


def test_fov_mask():
    m = G.fov_mask(512)
    assert m.shape == (512, 512) and np.array_equal(m, m.T)
    frac = m.mean()
    assert abs(frac - math.pi / 4 * 0.99) < 5e-3


def test_hd_config_mirror():
    cfg = G.hd_config(984, 888, 64, 512, 512, 64)
    assert cfg["acq"]["col_width"] == 1.0239 and abs(cfg["acq"]["col_height"] - 1.096439) < 1e-12
    assert cfg["recon"]["col_center"] == (888 + 1) / 2


def test_shepp_logan_range_and_seed():
    cfg = G.config_small()
    g = G.dd3_geometry(cfg)
    a = P.shepp_logan(g, seed=3)
    b = P.shepp_logan(g, seed=3)
    c = P.shepp_logan(g, seed=4)
    assert a.shape == g.vol_shape and a.dtype == np.float32
    assert np.array_equal(a, b) and not np.array_equal(a, c)
    assert a.min() >= 0 and a.max() <= 1 and a.max() > 0.9


def test_ellipsoid_mixture_range():
    cfg = G.helical_config(sid=541, sdd=949, DNU=64, DNV=8, col_width=2, col_height=2,
                           views_per_turn=10, turns=1, feed_mm_per_turn=1, N_xy=64, N_z=32,
                           voxel_xy=2.8, voxel_z=2.8)
    g = G.dd3_geometry(cfg)
    v = P.ellipsoid_mixture(g, seed=1)
    assert v.min() >= -0.02 and v.max() <= 2.02 and (v > 0).mean() > 0.05

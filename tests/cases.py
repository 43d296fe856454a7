"""Geometries shared by the oracle and GPU parity tests."""

import numpy as np

from ctrecon_amd import geometry as G
from ctrecon_amd import phantom as P


def small():
    """The small plain-method case (64^3, 256x16 detector, 720 views)."""
    cfg = G.config_small()
    return cfg, G.dd3_geometry(cfg, mask=G.fov_mask(64))


def tiny():
    cfg = G.helical_config(sid=60, sdd=100, DNU=12, DNV=4, col_width=4.0, col_height=3.0,
                           views_per_turn=8, turns=2, feed_mm_per_turn=6.0, N_xy=8, N_z=8,
                           voxel_xy=4.0, voxel_z=3.0)
    return cfg, G.dd3_geometry(cfg)


def ragged():
    """Sizes that are not multiples of anything, an off-centre volume and a mask with
    values 0, 1 and 2 (the forward projector multiplies by the byte value, the back
    projector keeps only voxels whose byte is 1)."""
    cfg = G.helical_config(sid=300, sdd=520, DNU=31, DNV=5, col_width=2.1, col_height=1.7,
                           views_per_turn=40, turns=1.5, feed_mm_per_turn=5.0, N_xy=23, N_z=37,
                           voxel_xy=1.9, voxel_z=1.1)
    m = np.ones((23, 23), dtype=np.uint8)
    rng = np.random.default_rng(7)
    m[rng.uniform(size=m.shape) < 0.15] = 0
    m[rng.uniform(size=m.shape) < 0.1] = 2
    g = G.dd3_geometry(cfg, mask=m)
    g.imgXCenter, g.imgYCenter, g.imgZCenter = 3.1, -2.2, 1.3
    return cfg, g


def flat():
    cfg = G.helical_config(sid=200, sdd=380, DNU=48, DNV=12, col_width=1.5, col_height=1.5,
                           views_per_turn=30, turns=2, feed_mm_per_turn=9.0, N_xy=32, N_z=24,
                           voxel_xy=1.5, voxel_z=1.5)
    cfg["recon"]["flat_det"] = 1
    return cfg, G.dd3_geometry(cfg, mask=G.fov_mask(32))


def clinical():
    cfg = G.config_clinical()
    return cfg, G.dd3_geometry(cfg, mask=G.fov_mask(512))


CASES = {"tiny": tiny, "ragged": ragged, "flat": flat, "small": small}


def volume(g, seed=0):
    if g.XN >= 32:
        return P.shepp_logan(g, seed=seed)
    return np.random.default_rng(seed).uniform(0, 1, g.vol_shape).astype(np.float32)


def sinogram(g, seed=1):
    return np.random.default_rng(seed).uniform(0, 1, g.prj_shape).astype(np.float32)

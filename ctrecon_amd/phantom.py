"""Seeded synthetic phantoms in the GPU volume layout ``vol[y, x, z]``.

* :func:`shepp_logan` -- the 3-D modified Shepp-Logan head (10 ellipsoids,
  densities 0..1) on the cube that the volume spans, with each centre jittered
  uniformly by up to ``jitter_mm`` from a seed.  The ellipsoid table is the
  published one (Toft, "The Radon Transform", pp. 199-200; the same table the
  reference's ``phantom3d('modified shepp-logan')`` uses, iterativeRecon/phantom3d.m:191-200).
* :func:`ellipsoid_mixture` -- random ellipsoids inside the FOV cylinder, painted
  in order (later ones overwrite), with a uniform-noise texture inside the support.

Voxel (x, y, z) has its centre at ((x-ocx)*dx, (y-ocy)*dx, (z-ocz)*dz) mm with
oc = ((N-1)/2 - centre/d), as in iterativeRecon/DD3_GPU.cu:1029-1031.
"""

from __future__ import annotations

import math

import numpy as np

from .geometry import Geometry

__all__ = ["shepp_logan", "ellipsoid_mixture", "voxel_centers"]

# A, a, b, c, x0, y0, z0, phi, theta, psi (angles in degrees), unit cube [-1, 1]^3
_MODIFIED_SHEPP_LOGAN = np.array([
    [1.0, 0.6900, 0.920, 0.810, 0.00, 0.0000, 0.00, 0.0, 0.0, 0.0],
    [-0.8, 0.6624, 0.874, 0.780, 0.00, -0.0184, 0.00, 0.0, 0.0, 0.0],
    [-0.2, 0.1100, 0.310, 0.220, 0.22, 0.0000, 0.00, -18.0, 0.0, 10.0],
    [-0.2, 0.1600, 0.410, 0.280, -0.22, 0.0000, 0.00, 18.0, 0.0, 10.0],
    [0.1, 0.2100, 0.250, 0.410, 0.00, 0.3500, -0.15, 0.0, 0.0, 0.0],
    [0.1, 0.0460, 0.046, 0.050, 0.00, 0.1000, 0.25, 0.0, 0.0, 0.0],
    [0.1, 0.0460, 0.046, 0.050, 0.00, -0.1000, 0.25, 0.0, 0.0, 0.0],
    [0.1, 0.0460, 0.023, 0.050, -0.08, -0.6050, 0.00, 0.0, 0.0, 0.0],
    [0.1, 0.0230, 0.023, 0.020, 0.00, -0.6060, 0.00, 0.0, 0.0, 0.0],
    [0.1, 0.0230, 0.046, 0.020, 0.06, -0.6050, 0.00, 0.0, 0.0, 0.0],
])


def voxel_centers(g: Geometry):
    """1-D voxel-centre coordinates (mm) along x, y, z."""
    ocx = (g.XN - 1.0) * 0.5 - g.imgXCenter / g.dx
    ocy = (g.YN - 1.0) * 0.5 - g.imgYCenter / g.dx
    ocz = (g.ZN - 1.0) * 0.5 - g.imgZCenter / g.dz
    xs = (np.arange(g.XN) - ocx) * g.dx
    ys = (np.arange(g.YN) - ocy) * g.dx
    zs = (np.arange(g.ZN) - ocz) * g.dz
    return xs, ys, zs


def _euler(phi, theta, psi):
    """z-x-z Euler rotation (angles in radians); rows map world -> ellipsoid frame."""
    cf, sf = math.cos(phi), math.sin(phi)
    ct, st = math.cos(theta), math.sin(theta)
    cp, sp = math.cos(psi), math.sin(psi)
    return np.array([
        [cp * cf - ct * sf * sp, cp * sf + ct * cf * sp, sp * st],
        [-sp * cf - ct * sf * cp, -sp * sf + ct * cf * cp, cp * st],
        [st * sf, -st * cf, ct],
    ])


def _paint(vol, xs, ys, zs, center, axes, R, value, mode):
    """Add (mode 'add') or overwrite (mode 'set') ``value`` inside one ellipsoid.
    Works on the bounding box only."""
    rad = float(max(axes))
    sel = []
    for arr, c in ((ys, center[1]), (xs, center[0]), (zs, center[2])):
        idx = np.nonzero(np.abs(arr - c) <= rad)[0]
        if idx.size == 0:
            return None
        sel.append(slice(idx[0], idx[-1] + 1))
    Y, X, Z = np.meshgrid(ys[sel[0]] - center[1], xs[sel[1]] - center[0],
                          zs[sel[2]] - center[2], indexing="ij")
    P = np.stack([X, Y, Z], axis=-1) @ R.T
    inside = ((P[..., 0] / axes[0]) ** 2 + (P[..., 1] / axes[1]) ** 2
              + (P[..., 2] / axes[2]) ** 2) <= 1.0
    sub = vol[sel[0], sel[1], sel[2]]
    if mode == "add":
        sub[inside] += value
    else:
        sub[inside] = value
    return sel, inside


def shepp_logan(g: Geometry, seed: int = 0, jitter_mm: float = 2.0) -> np.ndarray:
    """Modified 3-D Shepp-Logan on the cube the volume spans; float32 vol[y, x, z]."""
    xs, ys, zs = voxel_centers(g)
    half = np.array([g.XN * g.dx / 2, g.YN * g.dx / 2, g.ZN * g.dz / 2])
    rng = np.random.default_rng(seed)
    vol = np.zeros(g.vol_shape, dtype=np.float64)
    jit = rng.uniform(-jitter_mm, jitter_mm, (len(_MODIFIED_SHEPP_LOGAN), 3))
    jit[1] = jit[0]  # the two skull ellipsoids share one centre
    for row, dj in zip(_MODIFIED_SHEPP_LOGAN, jit):
        A, a, b, c, x0, y0, z0, phi, th, psi = row
        center = np.array([x0, y0, z0]) * half + dj
        axes = np.array([a, b, c]) * half
        R = _euler(math.radians(phi), math.radians(th), math.radians(psi))
        # phantom3d rotates coordinates about the cube centre, then offsets
        _paint_rotated_about_origin(vol, xs, ys, zs, center, axes, R, A, half)
    return np.clip(vol, 0.0, 1.0).astype(np.float32)


def _paint_rotated_about_origin(vol, xs, ys, zs, center, axes, R, value, half):
    # The table places the ellipsoid centre in the rotated frame: p' = R p, test
    # (p' - c)^2 / axes^2 <= 1.  Equivalent world-frame centre: R^T c.
    wc = R.T @ center
    _paint(vol, xs, ys, zs, wc, axes, R, value, "add")


def ellipsoid_mixture(g: Geometry, seed: int = 1, n: int = 40, density=(0.0, 2.0),
                      axes_xy=(5.0, 60.0), axes_z=(5.0, 40.0), texture_sigma=0.01,
                      r_frac=0.8) -> np.ndarray:
    """Random ellipsoid mixture: centres uniform in the FOV cylinder (r <= r_frac*R)
    and in z, semi-axes U(axes_xy) in-plane and U(axes_z) in z, rotation U(0, pi)
    about z, densities U(density) painted in order, plus a uniform texture of
    standard deviation ``texture_sigma`` inside the support.  float32 vol[y, x, z]."""
    xs, ys, zs = voxel_centers(g)
    rng = np.random.default_rng(seed)
    R_fov = 0.5 * min(g.XN, g.YN) * g.dx
    vol = np.zeros(g.vol_shape, dtype=np.float32)
    support = np.zeros(g.vol_shape, dtype=bool)
    zlo, zhi = zs.min(), zs.max()
    for _ in range(n):
        r = r_frac * R_fov * math.sqrt(rng.uniform())
        t = rng.uniform(0, 2 * math.pi)
        center = np.array([r * math.cos(t), r * math.sin(t), rng.uniform(zlo, zhi)])
        axes = np.array([rng.uniform(*axes_xy), rng.uniform(*axes_xy), rng.uniform(*axes_z)])
        rot = rng.uniform(0, math.pi)
        R = _euler(rot, 0.0, 0.0)
        val = rng.uniform(*density)
        out = _paint(vol, xs, ys, zs, center, axes, R, val, "set")
        if out is not None:
            sel, inside = out
            support[sel[0], sel[1], sel[2]] |= inside
    if texture_sigma > 0:
        h = texture_sigma * math.sqrt(3.0)
        noise = rng.uniform(-h, h, size=g.vol_shape).astype(np.float32)
        vol[support] += noise[support]
    return vol

"""CPU tests that pin the FP64 oracle (no GPU).

The reference ships no golden vectors for this path, so the oracle is pinned by
closed forms, by independent restatements (direct footprint overlap vs integral
images, a supersampled numerical footprint integral, dense-matrix restatements of
the loops) and by invariants (adjointness, linearity, mask semantics).
"""

import math

import numpy as np
import pytest

from conftest import rel_l2
from cases import CASES, ragged, sinogram, small, tiny, volume

from ctrecon_amd import geometry as G


# ---------------------------------------------------------------- boundaries
def test_boundaries_closed_form(oracle):
    # utilities.cuh:340-361: end cells extrapolated, interior midpoints
    b = oracle.boundaries([0.0, 1.0, 3.0, 6.0])
    assert np.allclose(b, [-0.5, 0.5, 2.0, 4.5, 7.5])
    b1 = oracle.boundaries([2.0])
    assert np.allclose(b1, [1.5, 2.5])


def test_det_paras_symmetric_detector(oracle):
    """calDetParas on a centred, symmetric detector: the centre indices are the middle
    of the array, the row pitch is col_height and dbeta is the cell angle."""
    cfg, g = small()
    cU, cV, dbeta, ddv = oracle.det_paras(g)
    assert abs(cU - (g.DNU - 1) / 2) < 1e-3
    assert abs(cV - (g.DNV - 1) / 2) < 1e-6
    assert abs(ddv - cfg["acq"]["col_height"]) < 1e-6
    assert abs(dbeta - cfg["acq"]["col_width"] / cfg["acq"]["sdd"]) < 1e-6


# ---------------------------------------------------------------- forward
@pytest.mark.parametrize("name", ["tiny", "ragged", "flat"])
def test_fp_sat_equals_direct(name, oracle):
    """Integral-image transcription == direct overlap (independent restatement)."""
    cfg, g = CASES[name]()
    vol = volume(g)
    a = oracle.fp(g, vol, "sat")
    b = oracle.fp(g, vol, "direct")
    assert rel_l2(a, b) < 1e-11


def test_fp_sat_equals_direct_small_subset(oracle):
    cfg, g = small()
    gs = g.subset(range(0, g.PN, 37))
    vol = volume(g)
    assert rel_l2(oracle.fp(gs, vol, "sat"), oracle.fp(gs, vol, "direct")) < 1e-11


def _ray_geometry(g, p, u, v):
    """Source, detector-cell centre and the cell's four boundary points (rotated)."""
    th = float(g.angles[p]) % (2 * math.pi)
    c, s = math.cos(th), math.sin(th)
    rot = lambda x, y: np.array([float(x) * c - float(y) * s, float(x) * s + float(y) * c])
    bx = oracle_boundaries_np(g.xds)
    by = oracle_boundaries_np(g.yds)
    bz = oracle_boundaries_np(g.zds)
    zp = float(g.zshifts[p])
    src = np.array([*rot(g.x0, g.y0), g.z0 + zp])
    D = np.array([*rot(g.xds[u], g.yds[u]), float(g.zds[v]) + zp])
    L = rot(bx[u], by[u])
    R = rot(bx[u + 1], by[u + 1])
    zlo, zhi = bz[v] + zp, bz[v + 1] + zp
    return src, D, L, R, zlo, zhi, abs(c) <= abs(s)


def oracle_boundaries_np(c):
    c = np.asarray(c, np.float64)
    b = np.empty(c.size + 1)
    b[0] = 1.5 * c[0] - 0.5 * c[1]
    b[1:-1] = 0.5 * (c[:-1] + c[1:])
    b[-1] = 1.5 * c[-1] - 0.5 * c[-2]
    return b


def _supersampled_ray(g, vol, p, u, v, n=48):
    """Independent numerical restatement of one ray sum: intersect the planes through
    the source and the detector cell's edges with every slice plane, sample the
    footprint rectangle on an n x n grid of points, look each point up in the
    piecewise-constant (masked) volume, and add (path length per slice) * mean."""
    src, D, L, R, zlo, zhi, xmarch = _ray_geometry(g, p, u, v)
    ocx = (g.XN - 1) / 2 - g.imgXCenter / g.dx
    ocy = (g.YN - 1) / 2 - g.imgYCenter / g.dx
    ocz = (g.ZN - 1) / 2 - g.imgZCenter / g.dz
    m = g.mask.reshape(g.YN, g.XN).astype(np.float64)
    f = vol.astype(np.float64) * m[:, :, None]
    a, b_ = (0, 1) if xmarch else (1, 0)  # marching axis, in-plane axis
    d = D - src
    K = g.dx * np.linalg.norm(d) / abs(d[a])
    NS = g.XN if xmarch else g.YN
    oc_a, oc_b = (ocx, ocy) if xmarch else (ocy, ocx)
    t = (np.arange(n) + 0.5) / n
    total = 0.0
    for ii in range(NS):
        X = (ii - oc_a) * g.dx
        lam = lambda P: (X - src[a]) / (P[a] - src[a])  # line src->P meets plane
        y1 = src[b_] + lam(L) * (L[b_] - src[b_])
        y2 = src[b_] + lam(R) * (R[b_] - src[b_])
        z1 = src[2] + lam(D) * (zlo - src[2])
        z2 = src[2] + lam(D) * (zhi - src[2])
        ys = y1 + (y2 - y1) * t
        zs = z1 + (z2 - z1) * t
        jj = np.floor(ys / g.dx + oc_b + 0.5).astype(int)
        kk = np.floor(zs / g.dz + ocz + 0.5).astype(int)
        NI = g.YN if xmarch else g.XN
        ok_j = (jj >= 0) & (jj < NI)
        ok_k = (kk >= 0) & (kk < g.ZN)
        if not ok_j.any() or not ok_k.any():
            continue
        jj_c, kk_c = np.clip(jj, 0, NI - 1), np.clip(kk, 0, g.ZN - 1)
        if xmarch:
            vals = f[jj_c[:, None], ii, kk_c[None, :]]
        else:
            vals = f[ii, jj_c[:, None], kk_c[None, :]]
        vals = vals * ok_j[:, None] * ok_k[None, :]
        total += vals.mean()
    return total * K


def test_fp_against_supersampled_integral(oracle):
    cfg, g = ragged()
    vol = volume(g, seed=4)
    ref = oracle.fp(g, vol, "direct")
    rng = np.random.default_rng(0)
    for _ in range(12):
        p, u, v = rng.integers(g.PN), rng.integers(g.DNU), rng.integers(g.DNV)
        want = _supersampled_ray(g, vol, p, u, v)
        got = ref[p, u, v]
        assert abs(got - want) <= 0.03 * max(abs(want), 1e-3) + 1e-3, (p, u, v, got, want)


def test_fp_uniform_slab_path_length(oracle):
    """All-ones volume and mask: a ray whose footprint stays inside the volume in every
    slice sums to (number of slices) * dx / |dir_axis| -- the chord through the slab.
    XN != YN so that the oblique views (30 deg off an axis) stay inside the box."""
    cfg = G.helical_config(sid=400, sdd=700, DNU=9, DNV=3, col_width=0.5, col_height=0.5,
                           views_per_turn=12, turns=1, feed_mm_per_turn=1.0, N_xy=40, N_z=60,
                           voxel_xy=1.0, voxel_z=1.0)
    g = G.dd3_geometry(cfg).subset([0, 1, 2, 3, 4])
    g.YN = 60
    g.mask = np.ones(g.XN * g.YN, np.uint8)
    vol = np.ones(g.vol_shape, np.float32)
    prj = oracle.fp(g, vol, "sat")
    for p in range(g.PN):
        src, D, L, R, zlo, zhi, xmarch = _ray_geometry(g, p, 4, 1)
        d = D - src
        a = 0 if xmarch else 1
        NS = g.XN if xmarch else g.YN
        want = NS * g.dx * np.linalg.norm(d) / abs(d[a])
        assert abs(prj[p, 4, 1] - want) < 1e-9 * want, (p, xmarch, prj[p, 4, 1], want)


def test_fp_chord_of_ball_sanity(oracle):
    """Voxelised ball vs the closed-form chord 2 sqrt(r^2 - q^2) of the central ray
    (the textbook chord formula of the reference's analytic projector,
    CreateProjSlice.cpp:121-134); discretisation-limited, so a loose bound."""
    cfg = G.helical_config(sid=300, sdd=500, DNU=11, DNV=3, col_width=0.5, col_height=0.5,
                           views_per_turn=4, turns=1, feed_mm_per_turn=0.0, N_xy=96, N_z=96,
                           voxel_xy=0.5, voxel_z=0.5)
    g = G.dd3_geometry(cfg).subset([0, 1])
    xs = (np.arange(96) - 47.5) * 0.5
    Y, X, Z = np.meshgrid(xs, xs, xs, indexing="ij")
    r = 18.0
    vol = ((X ** 2 + Y ** 2 + Z ** 2) <= r * r).astype(np.float32)
    prj = oracle.fp(g, vol, "sat")
    assert abs(prj[0, 5, 1] - 2 * r) / (2 * r) < 0.02
    assert abs(prj[1, 5, 1] - 2 * r) / (2 * r) < 0.02


def test_fp_linearity_and_mask_byte(oracle):
    cfg, g = ragged()
    rng = np.random.default_rng(1)
    # dyadic values so that 2a - 3b is exact in FP32
    a = (rng.integers(0, 1024, g.vol_shape) / 1024).astype(np.float32)
    b = (rng.integers(0, 1024, g.vol_shape) / 1024).astype(np.float32)
    fa, fb = oracle.fp(g, a, "sat"), oracle.fp(g, b, "sat")
    fab = oracle.fp(g, (2 * a - 3 * b).astype(np.float32), "sat")
    assert rel_l2(fab, 2 * fa - 3 * fb) < 1e-10
    # the reference multiplies the volume by the mask byte (DD3_GPU.cu:1004-1011)
    g1 = tiny()[1]
    g2 = tiny()[1]
    g2.mask = (g1.mask * 2).astype(np.uint8)
    v = volume(g1)
    assert rel_l2(oracle.fp(g2, v, "sat"), 2 * oracle.fp(g1, v, "sat")) < 1e-12


# ---------------------------------------------------------------- back
@pytest.mark.parametrize("name", ["tiny", "ragged", "flat"])
def test_bp_sat_equals_direct(name, oracle):
    cfg, g = CASES[name]()
    y = sinogram(g)
    assert rel_l2(oracle.bp(g, y, "sat"), oracle.bp(g, y, "direct")) < 1e-11


def test_bp_mask_semantics(oracle):
    """Only voxels whose mask byte is exactly 1 are back-projected (DD3_GPU.cu:1417)."""
    cfg, g = ragged()
    y = sinogram(g)
    out = oracle.bp(g, y, "sat")
    m = g.mask.reshape(g.YN, g.XN)
    assert np.all(out[m != 1] == 0)
    assert np.all(np.abs(out[m == 1]).sum(axis=-1) > 0)


def test_bp_reference_close_to_adjoint(oracle):
    """The reference back projector is not the exact adjoint, but it must agree with the
    transpose of the forward projector to a small tolerance; a sign, axis or index
    slip in either would not."""
    cfg, g = small()
    gs = g.subset(range(0, g.PN, 10))
    y = sinogram(gs)
    ref = oracle.bp(gs, y, "direct")
    adj = oracle.bp(gs, y, "adjoint")
    assert rel_l2(ref, adj) < 1e-3


def test_bp_ones_closed_form(oracle):
    """bp of an all-ones sinogram at a voxel whose shadow lies inside the detector is
    sum_p |cosVal| * shadow area (the SAT corners then difference to the area)."""
    cfg, g = small()
    gs = g.subset([350, 360, 450])  # views whose cone holds the voxel (zshift near 0)
    y = np.ones(gs.prj_shape, np.float32)
    out = oracle.bp(gs, y, "sat")
    cU, cV, dbeta, ddv = oracle.det_paras(gs)
    x, yy, z = 32, 32, 32
    ocx, ocy, ocz = (g.XN - 1) / 2, (g.YN - 1) / 2, (g.ZN - 1) / 2
    P = np.array([(x - ocx) * g.dx, (yy - ocy) * g.dx, (z - ocz) * g.dz])
    S2D = math.hypot(g.xds[0] - g.x0, g.yds[0] - g.y0)
    want = 0.0
    for p in range(gs.PN):
        th = float(gs.angles[p]) % (2 * math.pi)
        c, s = math.cos(th), math.sin(th)
        src = np.array([g.x0 * c - g.y0 * s, g.x0 * s + g.y0 * c, g.z0 + gs.zshifts[p]])
        d = P - src
        l2 = d[0] ** 2 + d[1] ** 2
        w = d[0] if abs(src[0]) > abs(src[1]) else d[1]
        area_u = abs(w / l2 * g.dx / dbeta)
        area_v = abs(g.dz * (S2D / ddv) / math.sqrt(l2))
        cos_val = abs(g.dx / w * math.sqrt(l2 + d[2] ** 2))
        want += cos_val * area_u * area_v
    assert abs(out[yy, x, z] - want) < 1e-9 * want


def test_matched_adjoint_identity(oracle):
    cfg, g = ragged()
    rng = np.random.default_rng(2)
    x = rng.uniform(0, 1, g.vol_shape).astype(np.float32)
    y = rng.uniform(0, 1, g.prj_shape).astype(np.float32)
    lhs = np.vdot(oracle.fp(g, x, "direct"), y.astype(np.float64))
    rhs = np.vdot(x.astype(np.float64), oracle.bp(g, y, "adjoint"))
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs)


# ---------------------------------------------------------------- loops
def _dense(oracle, g):
    """Dense A (forward) and B (reference back projector) of a tiny geometry, built
    column by column from the oracle."""
    nv, ns = int(np.prod(g.vol_shape)), int(np.prod(g.prj_shape))
    A = np.zeros((ns, nv))
    for i in range(nv):
        e = np.zeros(nv, np.float32)
        e[i] = 1
        A[:, i] = oracle.fp(g, e.reshape(g.vol_shape), "direct").ravel()
    B = np.zeros((nv, ns))
    for i in range(ns):
        e = np.zeros(ns, np.float32)
        e[i] = 1
        B[:, i] = oracle.bp(g, e.reshape(g.prj_shape), "direct").ravel()
    return A, B


def test_os_sart_matches_dense_restatement(oracle):
    cfg, g = tiny()
    A, B = _dense(oracle, g)
    xt = volume(g).astype(np.float64).ravel()
    b = (A @ xt).astype(np.float32)
    osNum, iters = 4, 2
    frame = g.DNU * g.DNV
    x = np.zeros_like(xt)
    for _ in range(iters):
        for s in range(osNum):
            views = np.arange(s, g.PN, osNum)
            rows = (views[:, None] * frame + np.arange(frame)[None, :]).ravel()
            As, Bs = A[rows], B[:, rows]
            rs = As @ np.ones(A.shape[1])
            cs = Bs @ np.ones(len(rows))
            tp = b.astype(np.float64).ravel()[rows] - As @ x
            tp = np.where(rs != 0, tp / np.where(rs != 0, rs, 1), tp)
            tm = Bs @ tp
            x = x + np.where(cs != 0, tm / np.where(cs != 0, cs, 1), tm)
    xo, _ = oracle.os_sart(g, b.reshape(g.prj_shape), osNum=osNum, iterNum=iters)
    assert rel_l2(xo.ravel(), x) < 1e-10


def test_conjgrad_matches_dense_restatement(oracle):
    cfg, g = tiny()
    A, B = _dense(oracle, g)
    xt = volume(g).astype(np.float64).ravel()
    b = (A @ xt).astype(np.float32).astype(np.float64)
    M = B @ A
    rhs = B @ b
    x = np.zeros_like(xt)
    r = rhs - M @ x
    p = r.copy()
    for _ in range(4):
        rr = r @ r
        Ap = M @ p
        alpha = rr / (p @ Ap)
        x = x + alpha * p
        r = r - alpha * Ap
        p = r + (r @ r) / rr * p
    xo, _ = oracle.conjgrad(g, b.reshape(g.prj_shape).astype(np.float32), iterNum=4)
    assert rel_l2(xo.ravel(), x) < 1e-8


def test_fp32_integral_images_lose_precision(oracle):
    """Documents why the GPU path integrates footprints directly: the FP32 integral-image
    restatement of the reference is orders of magnitude less accurate."""
    cfg, g = small()
    gs = g.subset(range(0, g.PN, 60))
    vol = volume(g)
    ref = oracle.fp(gs, vol, "sat")
    e32 = rel_l2(oracle.fp_f32(gs, vol), ref)
    assert 1e-6 < e32 < 1e-2

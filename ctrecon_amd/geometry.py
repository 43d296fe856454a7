"""Scan configuration and projector geometry.

``cfg`` mirrors the MATLAB ``cfg`` struct of the reference (``cfg.acq.*``,
``cfg.recon.*``) as nested dicts, and :func:`dd3_geometry` restates the
geometry block of ``dd3.m`` (iterativeRecon/dd3.m:86-186 plus the GPU axis swap
at :247-250 / :305-307): it turns a cfg into exactly the arrays the reference
passes to ``DD3_GPU('Proj'|'Back', ...)`` (iterativeRecon/dd3.m:251-258,
308-316).  MATLAB evaluates those expressions in ``single``; the same order of
single-precision operations is kept here with numpy float32.
"""

from __future__ import annotations

import copy
import math
from dataclasses import dataclass, field

import numpy as np

__all__ = [
    "Geometry",
    "hd_config",
    "helical_config",
    "config_small",
    "config_clinical",
    "dd3_geometry",
    "fov_mask",
    "view_subset",
]

f32 = np.float32


@dataclass
class Geometry:
    """Arguments of ``DD3Proj_gpu`` / ``DD3Back_gpu`` (iterativeRecon/DD3_GPU.cu:1132-1142,
    1766-1775) in the GPU coordinate system, minus the data pointers."""

    x0: float
    y0: float
    z0: float
    DNU: int
    DNV: int
    xds: np.ndarray  # float32 [DNU]
    yds: np.ndarray  # float32 [DNU]
    zds: np.ndarray  # float32 [DNV]
    imgXCenter: float
    imgYCenter: float
    imgZCenter: float
    angles: np.ndarray  # float32 [PN]  (already +pi/2, dd3.m:247)
    zshifts: np.ndarray  # float32 [PN]
    XN: int
    YN: int
    ZN: int
    dx: float
    dz: float
    mask: np.ndarray = field(default=None)  # uint8 [YN*XN] read as mask[y*XN + x]

    @property
    def PN(self) -> int:
        return int(self.angles.shape[0])

    @property
    def vol_shape(self):
        """GPU volume layout vol[y, x, z] (iterativeRecon/DD3_GPU.cu:44,1473)."""
        return (self.YN, self.XN, self.ZN)

    @property
    def prj_shape(self):
        """Sinogram layout prj[p, u, v] (iterativeRecon/DD3_GPU.cu:951)."""
        return (self.PN, self.DNU, self.DNV)

    def subset(self, views) -> "Geometry":
        g = copy.copy(self)
        idx = np.asarray(views, dtype=np.int64)
        g.angles = np.ascontiguousarray(self.angles[idx])
        g.zshifts = np.ascontiguousarray(self.zshifts[idx])
        return g


def hd_config(rotN, DNU, DNV, XN, YN, ZN):
    """Mirror of iterativeRecon/setHDConfig.m:1-25 (GE HD geometry)."""
    return {
        "acq": {
            "sid": 541.0,
            "sdd": 949.0,
            "views_per_rotation": rotN,
            "num_data_views": rotN,
            "row_count": DNV,
            "total_col_count": DNU,
            "col_width": 1.0239 / 888 * DNU,
            "col_height": 1.096439 / 64 * DNV,
            "rotation_direction": 1,
            "first_view_angle": 0.0,
            "first_view_zposition": 0.0,
            "helical_pitch": 0.0,
            "col_height_at_iso": 0.625,
        },
        "recon": {
            "col_center": (DNU + 1) / 2,
            "row_center": (DNV + 1) / 2,
            "recon_pixels_z": ZN,
            "recon_pixels_y": XN,
            "recon_pixels_x": YN,
            "dfov_mm": 250.0,
            "recon_slice_spacing": 0.625 / 64 * ZN,
            "recon_center_z": 0.0,
        },
    }


def helical_config(*, sid, sdd, DNU, DNV, col_width, col_height, views_per_turn, turns,
                   feed_mm_per_turn, N_xy, N_z, voxel_xy, voxel_z, col_height_at_iso=None,
                   centered_travel=True):
    """A helical cfg with a centred detector (``col_center = (DNU-1)/2``) and the
    table travel centred on z = 0 (``first_view_zposition = -travel/2``).

    ``helical_pitch`` is chosen so that the feed formula of dd3.m:178
    (``pitch * col_height_at_iso * row_count`` per turn) gives ``feed_mm_per_turn``.
    """
    if col_height_at_iso is None:
        col_height_at_iso = col_height * sid / sdd
    pitch = feed_mm_per_turn / (col_height_at_iso * DNV)
    num_views = int(views_per_turn * turns)
    travel = feed_mm_per_turn * turns * (num_views - 1) / num_views
    return {
        "acq": {
            "sid": float(sid),
            "sdd": float(sdd),
            "views_per_rotation": int(views_per_turn),
            "num_data_views": num_views,
            "row_count": int(DNV),
            "total_col_count": int(DNU),
            "col_width": float(col_width),
            "col_height": float(col_height),
            "rotation_direction": 1,
            "first_view_angle": 0.0,
            "first_view_zposition": -0.5 * travel if centered_travel else 0.0,
            "helical_pitch": float(pitch),
            "col_height_at_iso": float(col_height_at_iso),
        },
        "recon": {
            "col_center": (DNU - 1) / 2,
            "row_center": (DNV - 1) / 2,
            "recon_pixels_z": int(N_z),
            "recon_pixels_y": int(N_xy),
            "recon_pixels_x": int(N_xy),
            "dfov_mm": float(N_xy * voxel_xy),
            "recon_slice_spacing": float(voxel_z),
            "recon_center_z": 0.0,
        },
    }


def config_small():
    """The small plain-method case: 64^3 at 1 mm, 256x16 detector at 1.0 mm, 360 views/turn,
    2 turns, 15 mm/turn, sid 541 / sdd 949."""
    return helical_config(sid=541, sdd=949, DNU=256, DNV=16, col_width=1.0, col_height=1.0,
                          views_per_turn=360, turns=2, feed_mm_per_turn=15.0, N_xy=64, N_z=64,
                          voxel_xy=1.0, voxel_z=1.0)


def config_clinical():
    """The clinical-size case: 512x512x256 at 0.7 mm, 888x64 detector (1.0239 mm columns,
    1.096439 mm rows = 0.625 mm at iso), 984 views/turn, 4 turns, pitch 0.9."""
    return helical_config(sid=541, sdd=949, DNU=888, DNV=64, col_width=1.0239,
                          col_height=1.096439, views_per_turn=984, turns=4,
                          feed_mm_per_turn=0.9 * 64 * 0.625, N_xy=512, N_z=256, voxel_xy=0.7,
                          voxel_z=0.7, col_height_at_iso=0.625)


def config_training(N_xy=256, N_z=128):
    """The learned warm start's training scans: the clinical detector and trajectory
    (888x64, 984 views/turn, 4 turns, pitch 0.9) over a smaller grid at the same
    0.7 mm voxels (default 256x256x128: the central 179 mm x 90 mm)."""
    return helical_config(sid=541, sdd=949, DNU=888, DNV=64, col_width=1.0239,
                          col_height=1.096439, views_per_turn=984, turns=4,
                          feed_mm_per_turn=0.9 * 64 * 0.625, N_xy=N_xy, N_z=N_z, voxel_xy=0.7,
                          voxel_z=0.7, col_height_at_iso=0.625)


def fov_mask(n):
    """FOV mask of iterativeRecon/myDemo.m:40-48: ((i-0.5-N/2)/(N/2))^2 + ((j-0.5-N/2)/(N/2))^2
    < 0.99 for one-based i, j.  Returned as uint8 [n, n] (symmetric)."""
    i = np.arange(1, n + 1, dtype=np.float64)
    a = ((i - 0.5 - n / 2) / (n / 2)) ** 2
    return (a[:, None] + a[None, :] < 0.99).astype(np.uint8)


def view_subset(num_views, osIdx, osNum):
    """Zero-based views of subset ``osIdx`` (os_sart.m:15: viewSer(1+osIdx:osNum:end))."""
    return np.arange(osIdx, num_views, osNum, dtype=np.int64)


def dd3_geometry(cfg, view_ind=None, mask=None) -> Geometry:
    """Restates iterativeRecon/dd3.m:86-186 for the GPU path (+pi/2, swapped offsets).

    ``view_ind`` are zero-based view indices (dd3.m takes one-based ``view_ind``).
    ``mask`` is an [XN, YN] array (any integer type); it is stored the way the
    reference's ``int8([mask, mask'])`` is read by the GPU code (first XN*YN bytes,
    column-major, so mask[y*XN + x] = mask(x, y)).
    """
    acq, rec = cfg["acq"], cfg["recon"]
    x0 = f32(0.0)  # dd3.m:86,99 (shift_fs_mm = 0)
    if rec.get("para_mode"):
        y0 = f32(acq["sid"] * rec.get("para_scale", 1e10))  # dd3.m:114-120
    else:
        y0 = f32(acq["sid"])
    z0 = f32(0.0)
    nviews = int(acq["num_data_views"])
    # dd3.m:131: (0:N-1)*2*pi/single(vpr) + fva/180*pi, evaluated in single
    k = np.arange(nviews, dtype=np.float64)
    ang = (k * 2 * math.pi).astype(f32) / f32(acq["views_per_rotation"])
    ang = ang + f32(acq["first_view_angle"] / 180 * math.pi)
    if view_ind is None:
        view_ind = np.arange(nviews)
    ang = ang[np.asarray(view_ind, dtype=np.int64)]
    DNU, DNV = int(acq["total_col_count"]), int(acq["row_count"])
    sdd, sid = f32(acq["sdd"]), f32(acq["sid"])
    cols = np.arange(1, DNU + 1, dtype=f32)
    if not rec.get("flat_det"):
        # dd3.m:150-152
        alphas = (cols - f32(rec["col_center"] + 1)) * f32(acq["col_width"]) / sdd
        xds = np.sin(alphas) * sdd
        yds = sid - sdd * np.cos(alphas)
    else:
        # dd3.m:156-157
        xds = (cols - f32(rec["col_center"] + 1)) * f32(acq["col_width"])
        yds = np.full(DNU, f32(acq["sid"] - acq["sdd"]), dtype=f32)
    rows = np.arange(1, DNV + 1, dtype=f32)
    zds = (rows - f32(DNV + 1) * f32(0.5)) * f32(acq["col_height"])  # dd3.m:159
    cen_x, cen_y = 0.0, 0.0  # dd3.m:168-173 (no targcen)
    zoffset = f32(rec["recon_center_z"])
    # dd3.m:178
    zsh = ((ang - f32(acq["first_view_angle"] / 180 * math.pi)) / f32(2) / f32(math.pi)
           * f32(acq["helical_pitch"]) * f32(acq["col_height_at_iso"]) * f32(DNV)
           + f32(acq["first_view_zposition"]))
    dx = f32(rec["dfov_mm"]) / f32(max(rec["recon_pixels_x"], rec["recon_pixels_y"]))  # :180
    XN, YN = int(rec["recon_pixels_x"]), int(rec["recon_pixels_y"])
    if mask is None:
        m = np.ones((YN, XN), dtype=np.uint8)
    else:
        m = np.asarray(mask).astype(np.int8).view(np.uint8)
        m = np.ascontiguousarray(m.T)  # column-major (x fastest) byte order
    return Geometry(
        x0=float(x0), y0=float(y0), z0=float(z0), DNU=DNU, DNV=DNV,
        xds=np.ascontiguousarray(xds, dtype=f32), yds=np.ascontiguousarray(yds, dtype=f32),
        zds=np.ascontiguousarray(zds, dtype=f32),
        imgXCenter=float(f32(-cen_y)), imgYCenter=float(f32(cen_x)), imgZCenter=float(zoffset),
        angles=np.ascontiguousarray(ang + f32(math.pi / 2), dtype=f32),  # dd3.m:247
        zshifts=np.ascontiguousarray(zsh, dtype=f32),
        XN=XN, YN=YN, ZN=int(rec["recon_pixels_z"]),
        dx=float(dx), dz=float(f32(rec["recon_slice_spacing"])),
        mask=np.ascontiguousarray(m.reshape(-1), dtype=np.uint8),
    )

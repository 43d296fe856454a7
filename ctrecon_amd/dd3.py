"""MATLAB-signature front end: ``dd3``, ``os_sart`` and ``conjgrad``.

These mirror the reference's MATLAB functions argument for argument
(iterativeRecon/dd3.m:68, iterativeRecon/os_sart.m:1, iterativeRecon/conjgrad.m:1)
so that a caller of the reference finds the same surface.  Arrays use MATLAB's
shapes and column-major (Fortran) memory order:

* image    ``[recon_pixels_z, recon_pixels_x, recon_pixels_y]`` (z fastest)
* sinogram ``[row_count, total_col_count, nviews]`` (detector row fastest)

``view_ind`` is zero-based (MATLAB's is one-based).  The projector strings of
dd3.m:233-246 / :291-304 select prjMode as there, and prjMode selects the model as
DD3Proj_gpu / DD3Back_gpu do (DD3_GPU.cu:1144-1183, 1778-1815): pseudo distance-driven
for ``fp_gpu_volume_rendering`` / ``fp_gpu_pseudo_dd`` and ``bp_gpu_pseudo_dd`` /
``bp_gpu_soft_interp``, the FP64 branchless forward model for ``fp_gpu_soft_interp``,
the z-line back projector for ``bp_gpu_zline``, the branchless model otherwise.
``bp_gpu_matched`` (an addition of this library) selects the exact adjoint of the
branchless forward projector.
"""

from __future__ import annotations

import ctypes

import numpy as np

from ._lib import CtrError, CTR_ERR_ARG, CTR_ERR_UNSUPPORTED, check, lib
from .geometry import dd3_geometry
from .plan import DeviceArray, Plan

__all__ = ["dd3", "os_sart", "conjgrad", "projector_mode"]

_FP_MODES = {"fp_gpu_branchless": 0, "fp_gpu_volume_rendering": 1, "fp_gpu_soft_interp": 2,
             "fp_gpu_pseudo_dd": 3, "fp_gpu_branches": 4}
_BP_MODES = {"bp_gpu_branchless": 0, "bp_gpu_pseudo_dd": 1, "bp_gpu_soft_interp": 2,
             "bp_gpu_zline": 3, "bp_gpu_branches": 4, "bp_gpu_matched": 16}


def projector_mode(flag_fw: str) -> int:
    """prjMode of a projector string (dd3.m:233-246, 291-304; unknown -> 0)."""
    if flag_fw[:6] not in ("fp_gpu", "bp_gpu"):
        raise CtrError(CTR_ERR_UNSUPPORTED,
                       f"'{flag_fw}': only the GPU projectors are provided (the reference's "
                       "CPU DD3_mex is not part of it either)")
    table = _FP_MODES if flag_fw[:2] == "fp" else _BP_MODES
    return table.get(flag_fw, 0)


def _geo_argv(g):
    keep = [np.ascontiguousarray(a, dtype=np.float32)
            for a in (g.xds, g.yds, g.zds, g.angles, g.zshifts)]
    argv = [g.x0, g.y0, g.z0, g.DNU, g.DNV, keep[0].ctypes.data, keep[1].ctypes.data,
            keep[2].ctypes.data, g.imgXCenter, g.imgYCenter, g.imgZCenter, keep[3].ctypes.data,
            keep[4].ctypes.data, g.PN, g.XN, g.YN, g.ZN]
    return argv, keep


def dd3(flag_fw, cfg, in_, view_ind=None, mask=None, bp_type=None, gpu_id=0):
    """``[out] = dd3(flag_fw, cfg, in, view_ind, mask, bp_type)`` (iterativeRecon/dd3.m:68).

    fp: ``in`` is the image [Z, X, Y]; returns the sinogram [DNV, DNU, nviews].
    bp: ``in`` is the sinogram; returns the image [Z, X, Y].
    """
    rec, acq = cfg["recon"], cfg["acq"]
    kind = flag_fw[:2]
    if kind not in ("fp", "bp"):
        raise CtrError(CTR_ERR_ARG, "must specify fp or bp projector")
    img_shape = (rec["recon_pixels_z"], rec["recon_pixels_x"], rec["recon_pixels_y"])
    if view_ind is None:
        view_ind = np.arange(acq["num_data_views"])
    view_ind = np.asarray(view_ind, dtype=np.int64)
    prjMode = projector_mode(flag_fw)
    g = dd3_geometry(cfg, view_ind, mask)
    argv, keep = _geo_argv(g)
    m = np.ascontiguousarray(g.mask, dtype=np.uint8)
    if kind == "fp":
        img = np.asarray(in_, dtype=np.float32)
        if img.shape != img_shape:  # dd3.m:73-75
            raise CtrError(CTR_ERR_ARG, "Image dimensionality appears incorrect; it should be "
                                        "z-ordered.")
        gpu = _to_gpu_volume(img, g)  # permute [1 3 2] (dd3.m:248)
        prj = np.empty(g.PN * g.DNU * g.DNV, dtype=np.float32)
        check(lib().ctr_dd3_proj(*argv, gpu.ctypes.data, prj.ctypes.data, g.dx, g.dz,
                                 m.ctypes.data, gpu_id, prjMode))
        return prj.reshape(g.PN, g.DNU, g.DNV).transpose(2, 1, 0)  # [DNV, DNU, PN], F order
    sino = np.asarray(in_, dtype=np.float32)
    if sino.shape != (acq["row_count"], acq["total_col_count"], view_ind.size):  # dd3.m:137
        raise CtrError(CTR_ERR_ARG, "Sino dimensionality appears incorrect; it should be "
                                    "row-ordered.")
    flat = np.ascontiguousarray(sino.ravel(order="F"))
    vol = np.empty(g.XN * g.YN * g.ZN, dtype=np.float32)
    check(lib().ctr_dd3_back(*argv, vol.ctypes.data, flat.ctypes.data, g.dx, g.dz, m.ctypes.data,
                             gpu_id, int(bp_type or 0), prjMode))
    # kernel order [Z, X, Y] (Fortran); dd3.m:317 permutes [1 3 2] back
    out = vol.reshape(g.YN, g.XN, g.ZN).transpose(2, 1, 0)
    return np.transpose(out, (0, 2, 1))


def _to_gpu_volume(img, g):
    """MATLAB image [Z, X, Y] -> GPU layout vol[y, x, z] after dd3.m's permute [1 3 2]
    (column-major [Z, Y, X] is row-major [X, Y, Z])."""
    return np.ascontiguousarray(np.transpose(img, (0, 2, 1)).ravel(order="F")).reshape(
        g.YN, g.XN, g.ZN)


def _from_gpu_volume(vol, g):
    out = vol.reshape(g.YN, g.XN, g.ZN).transpose(2, 1, 0)
    return np.transpose(out, (0, 2, 1))


def _to_gpu_sino(sino):
    return np.ascontiguousarray(np.asarray(sino, np.float32).transpose(2, 1, 0))


def os_sart(initImg, prj, cfg, mask, osNum, iterNum, bp="ref", gpu_id=0, return_residuals=False):
    """``reconImg = os_sart(initImg, prj, cfg, mask, osNum, iterNum)``
    (iterativeRecon/os_sart.m:1-44), the whole loop on the GPU."""
    g = dd3_geometry(cfg, None, mask)
    plan = Plan(g, gpu_id)
    dx = DeviceArray.from_host(_to_gpu_volume(np.asarray(initImg, np.float32), g), gpu_id)
    dp = DeviceArray.from_host(_to_gpu_sino(prj), gpu_id)
    res = plan.os_sart(dx, dp, int(osNum), int(iterNum), bp)
    out = _from_gpu_volume(dx.get(), g)
    return (out, res) if return_residuals else out


def conjgrad(initImg, prj, cfg, mask, iterNum, bp="ref", gpu_id=0, return_residuals=False):
    """``reconImg = conjgrad(initImg, prj, cfg, mask, iterNum)``
    (iterativeRecon/conjgrad.m:1-35), the whole loop on the GPU with FP64 scalars."""
    g = dd3_geometry(cfg, None, mask)
    plan = Plan(g, gpu_id)
    dx = DeviceArray.from_host(_to_gpu_volume(np.asarray(initImg, np.float32), g), gpu_id)
    dp = DeviceArray.from_host(_to_gpu_sino(prj), gpu_id)
    res = plan.conjgrad(dx, dp, int(iterNum), bp)
    out = _from_gpu_volume(dx.get(), g)
    return (out, res) if return_residuals else out

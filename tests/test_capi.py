"""The C ABI library: loads, exports every declared symbol, validates arguments
(no GPU needed -- every case here fails before any device work)."""

import ctypes
import os
import re

import numpy as np
import pytest

from ctrecon_amd import _lib
from ctrecon_amd import geometry as G
from ctrecon_amd._lib import CtrError, CtrGeometry, lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    txt = open(os.path.join(ROOT, "include", "ctrecon.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(ctr_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    L = lib()
    names = _declared()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(L, n)]
    assert missing == []
    assert set(names) == set(_lib.EXPORTS)


def test_version_and_error_string():
    assert lib().ctr_version().startswith(b"ctrecon_amd")
    assert isinstance(lib().ctr_last_error(), bytes)


def _geometry_struct(g, keep):
    arrs = [np.ascontiguousarray(a, np.float32) for a in (g.xds, g.yds, g.zds, g.angles,
                                                          g.zshifts)]
    keep.extend(arrs)
    return CtrGeometry(x0=g.x0, y0=g.y0, z0=g.z0, DNU=g.DNU, DNV=g.DNV,
                       xds=arrs[0].ctypes.data, yds=arrs[1].ctypes.data,
                       zds=arrs[2].ctypes.data, imgXCenter=0, imgYCenter=0, imgZCenter=0,
                       angles=arrs[3].ctypes.data, zshifts=arrs[4].ctypes.data, PN=g.PN,
                       XN=g.XN, YN=g.YN, ZN=g.ZN, dx=g.dx, dz=g.dz, mask=None)


def _create(cg):
    h = ctypes.c_void_p()
    rc = lib().ctr_plan_create(ctypes.byref(cg) if cg is not None else None, 0, ctypes.byref(h))
    return rc, lib().ctr_last_error().decode()


def test_plan_create_rejects_bad_geometry():
    g = G.dd3_geometry(G.config_small())
    keep = []
    assert _create(None)[0] == _lib.CTR_ERR_ARG
    cg = _geometry_struct(g, keep)
    cg.DNU = 0
    rc, msg = _create(cg)
    assert rc == _lib.CTR_ERR_ARG and "sizes" in msg
    cg = _geometry_struct(g, keep)
    cg.dx = 0.0
    assert _create(cg)[0] == _lib.CTR_ERR_ARG
    bad = g.subset(range(g.PN))
    bad.angles = bad.angles.copy()
    bad.angles[3] = np.nan
    assert _create(_geometry_struct(bad, keep))[0] == _lib.CTR_ERR_NUMERIC
    dec = g.subset(range(g.PN))
    dec.zds = dec.zds[::-1].copy()
    rc, msg = _create(_geometry_struct(dec, keep))
    assert rc == _lib.CTR_ERR_UNSUPPORTED and "zds" in msg
    cg = _geometry_struct(g, keep)
    cg.xds = None
    assert _create(cg)[0] == _lib.CTR_ERR_ARG


def test_reference_entry_points_argument_checks():
    """Checked before any device work: the matched-adjoint bit is refused with the
    pseudo-DD back projector (it exists for the branchless model only), and NULL data
    buffers are argument errors."""
    g = G.dd3_geometry(G.config_small())
    z = np.zeros(4, np.float32)
    for mode in (1 | 16, 2 | 16):
        rc = lib().ctr_dd3_back(g.x0, g.y0, g.z0, g.DNU, g.DNV, z.ctypes.data, z.ctypes.data,
                                z.ctypes.data, 0, 0, 0, z.ctypes.data, z.ctypes.data, g.PN,
                                g.XN, g.YN, g.ZN, z.ctypes.data, z.ctypes.data, g.dx, g.dz,
                                None, 0, 0, mode)
        assert rc == _lib.CTR_ERR_UNSUPPORTED
    for mode in (0, 1, 2, 3):
        rc = lib().ctr_dd3_proj(g.x0, g.y0, g.z0, g.DNU, g.DNV, z.ctypes.data, z.ctypes.data,
                                z.ctypes.data, 0, 0, 0, z.ctypes.data, z.ctypes.data, g.PN,
                                g.XN, g.YN, g.ZN, None, None, g.dx, g.dz, None, 0, mode)
        assert rc == _lib.CTR_ERR_ARG
        rc = lib().ctr_dd3_back(g.x0, g.y0, g.z0, g.DNU, g.DNV, z.ctypes.data, z.ctypes.data,
                                z.ctypes.data, 0, 0, 0, z.ctypes.data, z.ctypes.data, g.PN,
                                g.XN, g.YN, g.ZN, None, None, g.dx, g.dz, None, 0, 0, mode)
        assert rc == _lib.CTR_ERR_ARG


def test_null_plan_calls_fail_cleanly():
    L = lib()
    assert L.ctr_plan_fp(None, None, None, 0, 1, 1, None) == _lib.CTR_ERR_ARG
    assert L.ctr_plan_bp(None, None, None, 0, 1, 1, 0, None) == _lib.CTR_ERR_ARG
    assert L.ctr_os_sart(None, None, None, 10, 1, 0, None) == _lib.CTR_ERR_ARG
    assert L.ctr_conjgrad(None, None, None, 1, 0, None) == _lib.CTR_ERR_ARG
    assert L.ctr_plan_destroy(None) == _lib.CTR_OK


def test_matlab_surface_validates_shapes():
    """dd3.m:73-75 and :137-139 raise on wrongly shaped images and sinograms."""
    from ctrecon_amd import dd3 as D

    cfg = G.config_small()
    with pytest.raises(CtrError):
        D.dd3("fp_gpu_branchless", cfg, np.zeros((64, 64, 63), np.float32))
    with pytest.raises(CtrError):
        D.dd3("bp_gpu_branchless", cfg, np.zeros((16, 256, 5), np.float32), [0, 1, 2])
    with pytest.raises(CtrError):
        D.dd3("xx", cfg, np.zeros((64, 64, 64), np.float32))
    with pytest.raises(CtrError):
        D.dd3("fp", cfg, np.zeros((64, 64, 64), np.float32))  # CPU DD3_mex: not provided
    assert D.projector_mode("fp_gpu_branchless") == 0
    assert D.projector_mode("bp_gpu_matched") == 16
    assert D.projector_mode("bp_gpu_zline") == 3


def test_null_handles_of_the_newer_entry_points():
    """Sessions, batches, the network, storage and mask entry points reject NULL
    handles and buffers with CTR_ERR_ARG (before any device work)."""
    L = lib()
    p = ctypes.c_void_p()
    assert L.ctr_plan_set_storage(None, 0) == _lib.CTR_ERR_ARG
    assert L.ctr_plan_mask_volume(None, None, None) == _lib.CTR_ERR_ARG
    assert L.ctr_plan_fp_bf16(None, None, None, 0, 1, 1, None) == _lib.CTR_ERR_ARG
    assert L.ctr_plan_bp_bf16(None, None, None, 0, 1, 1, 0, None) == _lib.CTR_ERR_ARG
    assert L.ctr_f32_to_bf16(None, None, 4, None) == _lib.CTR_ERR_ARG
    assert L.ctr_bf16_to_f32(None, None, 4, None) == _lib.CTR_ERR_ARG
    assert L.ctr_device_sub(None, None, None, 4, None) == _lib.CTR_ERR_ARG
    assert L.ctr_device_axpy(None, None, 1.0, 4, None) == _lib.CTR_ERR_ARG
    assert L.ctr_sart_create(None, None, 1, 0, ctypes.byref(p)) == _lib.CTR_ERR_ARG
    assert L.ctr_sart_create(None, None, 1, 0, None) == _lib.CTR_ERR_ARG
    assert L.ctr_sart_sweeps(None, None, 1, 0.0, None, None, None) == _lib.CTR_ERR_ARG
    assert L.ctr_sart_destroy(None) == _lib.CTR_OK
    assert L.ctr_batch_create(None, 0, None, None, 1, 0, ctypes.byref(p)) == _lib.CTR_ERR_ARG
    assert L.ctr_batch_sweeps(None, 1, 0.0, None, None, None) == _lib.CTR_ERR_ARG
    assert L.ctr_batch_set_graph(None, 1) == _lib.CTR_ERR_ARG
    assert L.ctr_batch_set_tuning(None, 1, 0, 0, 0, 0, 0) == _lib.CTR_ERR_ARG
    assert L.ctr_batch_get_info(None, None) == _lib.CTR_ERR_ARG
    assert L.ctr_batch_destroy(None) == _lib.CTR_OK
    assert L.ctr_unet_create(0, None) == _lib.CTR_ERR_ARG
    assert L.ctr_unet_set_params(None, None, 0) == _lib.CTR_ERR_ARG
    assert L.ctr_unet_get_params(None, None, 0) == _lib.CTR_ERR_ARG
    assert L.ctr_unet_set_scales(None, 1.0, 1.0, 1.0) == _lib.CTR_ERR_ARG
    assert L.ctr_unet_forward(None, None, None, None, 4, 4, 32, None) == _lib.CTR_ERR_ARG
    assert L.ctr_unet_loss_grad(None, None, None, None, 4, 4, 32, None, None) == _lib.CTR_ERR_ARG
    assert L.ctr_unet_train_step(None, None, None, None, 4, 4, 32, 1e-3, None) == _lib.CTR_ERR_ARG
    assert L.ctr_unet_destroy(None) == _lib.CTR_OK


def test_unet_parameter_count_matches_the_layer_table():
    from ctrecon_amd.unet import LAYERS, num_params

    assert lib().ctr_unet_num_params() == num_params()
    assert num_params() == sum(27 * ci * co + co for ci, co in LAYERS) + 17
    from oracle import unet_oracle as U

    assert U.LAYERS == LAYERS

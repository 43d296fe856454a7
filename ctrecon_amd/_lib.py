"""ctypes binding of libctrecon_amd.so (the C ABI declared in include/ctrecon.h).

The library is loaded from this package folder only.  There is no fallback: if
the shared object is missing or a call fails, a :class:`CtrError` is raised.
"""

from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libctrecon_amd.so")

CTR_OK = 0
CTR_ERR_ARG = 1
CTR_ERR_HIP = 2
CTR_ERR_UNSUPPORTED = 3
CTR_ERR_NOMEM = 4
CTR_ERR_NUMERIC = 5
CTR_BP_REFERENCE = 0
CTR_BP_MATCHED = 1
CTR_METHOD_DIRECT = 0
CTR_METHOD_SAT = 1
CTR_METHOD_PSEUDO_DD = 2
CTR_METHOD_ZLINE = 3
CTR_STORAGE_F32 = 0
CTR_STORAGE_BF16 = 1
CTR_STORAGE_BF16_SINO = 2

_NAMES = {0: "CTR_OK", 1: "CTR_ERR_ARG", 2: "CTR_ERR_HIP", 3: "CTR_ERR_UNSUPPORTED",
          4: "CTR_ERR_NOMEM", 5: "CTR_ERR_NUMERIC"}

# Every symbol include/ctrecon.h declares.
EXPORTS = [
    "ctr_dd3_proj", "ctr_dd3_back", "ctr_plan_create", "ctr_plan_destroy", "ctr_plan_set_method",
    "ctr_plan_set_storage", "ctr_plan_mask_volume", "ctr_plan_fp", "ctr_plan_fp_bf16", "ctr_plan_bp_bf16",
    "ctr_f32_to_bf16", "ctr_bf16_to_f32",
    "ctr_plan_bp", "ctr_os_sart", "ctr_os_sart_run", "ctr_cone_voxels", "ctr_conjgrad",
    "ctr_residual", "ctr_residual_noalloc", "ctr_device_count",
    "ctr_malloc", "ctr_free", "ctr_memcpy_h2d", "ctr_memcpy_d2h", "ctr_memset",
    "ctr_synchronize", "ctr_device_copy", "ctr_device_sub", "ctr_device_axpy", "ctr_last_error", "ctr_version",
    "ctr_sart_create", "ctr_sart_sweeps", "ctr_sart_setup_ms", "ctr_sart_destroy",
    "ctr_sart_set_graph", "ctr_sart_graph_info", "ctr_debug_fail_alloc_after",
    "ctr_batch_create", "ctr_batch_sweeps", "ctr_batch_set_graph", "ctr_batch_set_tuning",
    "ctr_batch_get_info", "ctr_batch_destroy",
    "ctr_unet_num_params", "ctr_unet_create", "ctr_unet_destroy", "ctr_unet_set_params",
    "ctr_unet_get_params", "ctr_unet_set_scales", "ctr_unet_forward", "ctr_unet_loss_grad",
    "ctr_unet_train_step",
]


class CtrError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{_NAMES.get(code, code)}: {msg}")
        self.code = code


class CtrGeometry(ctypes.Structure):
    _fields_ = [
        ("x0", ctypes.c_float), ("y0", ctypes.c_float), ("z0", ctypes.c_float),
        ("DNU", ctypes.c_int), ("DNV", ctypes.c_int),
        ("xds", ctypes.c_void_p), ("yds", ctypes.c_void_p), ("zds", ctypes.c_void_p),
        ("imgXCenter", ctypes.c_float), ("imgYCenter", ctypes.c_float),
        ("imgZCenter", ctypes.c_float),
        ("angles", ctypes.c_void_p), ("zshifts", ctypes.c_void_p), ("PN", ctypes.c_int),
        ("XN", ctypes.c_int), ("YN", ctypes.c_int), ("ZN", ctypes.c_int),
        ("dx", ctypes.c_float), ("dz", ctypes.c_float),
        ("mask", ctypes.c_void_p),
    ]


class CtrSartStats(ctypes.Structure):
    _fields_ = [
        ("sweeps", ctypes.c_int), ("fp_launches", ctypes.c_int), ("bp_launches", ctypes.c_int),
        ("setup_ms", ctypes.c_double), ("fp_ms", ctypes.c_double), ("bp_ms", ctypes.c_double),
    ]


class CtrBatchInfo(ctypes.Structure):
    _fields_ = [
        ("fp_colw", ctypes.c_int), ("fp_grid", ctypes.c_int), ("fp_strided", ctypes.c_int),
        ("bp_zg", ctypes.c_int), ("bp_grid", ctypes.c_int), ("bp_strided", ctypes.c_int),
        ("setup_ms", ctypes.c_double), ("tune_ms", ctypes.c_double),
        ("capture_ms", ctypes.c_double), ("fp_items", ctypes.c_longlong),
        ("bp_items", ctypes.c_longlong), ("from_cache", ctypes.c_int),
    ]


_lib = None


def lib():
    """Load libctrecon_amd.so; raise if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise CtrError(CTR_ERR_ARG, f"{LIB_PATH} is missing; run python -m ctrecon_amd.build")
        L = ctypes.CDLL(LIB_PATH)
        f, i, p, d = ctypes.c_float, ctypes.c_int, ctypes.c_void_p, ctypes.c_double
        geo = [f, f, f, i, i, p, p, p, f, f, f, p, p, i, i, i, i]
        L.ctr_dd3_proj.argtypes = geo + [p, p, f, f, p, i, i]
        L.ctr_dd3_back.argtypes = geo + [p, p, f, f, p, i, i, i]
        L.ctr_plan_create.argtypes = [ctypes.POINTER(CtrGeometry), i, ctypes.POINTER(p)]
        L.ctr_plan_destroy.argtypes = [p]
        L.ctr_plan_set_method.argtypes = [p, i]
        L.ctr_plan_set_storage.argtypes = [p, i]
        L.ctr_plan_mask_volume.argtypes = [p, p, p]
        L.ctr_plan_fp.argtypes = [p, p, p, i, i, i, p]
        L.ctr_plan_fp_bf16.argtypes = [p, p, p, i, i, i, p]
        L.ctr_plan_bp_bf16.argtypes = [p, p, p, i, i, i, i, p]
        L.ctr_f32_to_bf16.argtypes = [p, p, ctypes.c_ulonglong, p]
        L.ctr_bf16_to_f32.argtypes = [p, p, ctypes.c_ulonglong, p]
        L.ctr_plan_bp.argtypes = [p, p, p, i, i, i, i, p]
        L.ctr_os_sart.argtypes = [p, p, p, i, i, i, p]
        L.ctr_os_sart_run.argtypes = [p, p, p, i, i, i, d, p, p, ctypes.POINTER(CtrSartStats)]
        L.ctr_cone_voxels.argtypes = [p, i, i, i, p]
        L.ctr_conjgrad.argtypes = [p, p, p, i, i, p]
        L.ctr_residual.argtypes = [p, p, p, ctypes.POINTER(d)]
        L.ctr_residual_noalloc.argtypes = [p, p, p, ctypes.POINTER(d)]
        L.ctr_device_count.argtypes = [ctypes.POINTER(i)]
        L.ctr_malloc.argtypes = [ctypes.POINTER(p), ctypes.c_ulonglong, i]
        L.ctr_free.argtypes = [p]
        L.ctr_memcpy_h2d.argtypes = [p, p, ctypes.c_ulonglong]
        L.ctr_memcpy_d2h.argtypes = [p, p, ctypes.c_ulonglong]
        L.ctr_memset.argtypes = [p, i, ctypes.c_ulonglong]
        L.ctr_synchronize.argtypes = [i]
        L.ctr_device_copy.argtypes = [p, p, ctypes.c_ulonglong, p]
        L.ctr_device_sub.argtypes = [p, p, p, ctypes.c_ulonglong, p]
        L.ctr_device_axpy.argtypes = [p, p, d, ctypes.c_ulonglong, p]
        L.ctr_sart_create.argtypes = [p, p, i, i, ctypes.POINTER(p)]
        L.ctr_sart_sweeps.argtypes = [p, p, i, d, p, p, ctypes.POINTER(CtrSartStats)]
        L.ctr_sart_setup_ms.argtypes = [p]
        L.ctr_sart_destroy.argtypes = [p]
        L.ctr_sart_set_graph.argtypes = [p, i]
        L.ctr_sart_graph_info.argtypes = [p, ctypes.POINTER(i), ctypes.POINTER(d)]
        L.ctr_debug_fail_alloc_after.argtypes = [ctypes.c_longlong]
        L.ctr_batch_create.argtypes = [p, i, p, p, i, i, ctypes.POINTER(p)]
        L.ctr_batch_sweeps.argtypes = [p, i, d, p, p, ctypes.POINTER(i)]
        L.ctr_batch_set_graph.argtypes = [p, i]
        L.ctr_batch_set_tuning.argtypes = [p, i, i, i, i, i, i]
        L.ctr_batch_get_info.argtypes = [p, ctypes.POINTER(CtrBatchInfo)]
        L.ctr_batch_destroy.argtypes = [p]
        ll = ctypes.c_longlong
        L.ctr_unet_num_params.argtypes = []
        L.ctr_unet_create.argtypes = [i, ctypes.POINTER(p)]
        L.ctr_unet_destroy.argtypes = [p]
        L.ctr_unet_set_params.argtypes = [p, p, ll]
        L.ctr_unet_get_params.argtypes = [p, p, ll]
        L.ctr_unet_set_scales.argtypes = [p, f, f, f]
        L.ctr_unet_forward.argtypes = [p, p, p, p, i, i, i, p]
        L.ctr_unet_loss_grad.argtypes = [p, p, p, p, i, i, i, p, ctypes.POINTER(d)]
        L.ctr_unet_train_step.argtypes = [p, p, p, p, i, i, i, f, ctypes.POINTER(d)]
        L.ctr_last_error.restype = ctypes.c_char_p
        L.ctr_version.restype = ctypes.c_char_p
        for name in EXPORTS:
            fn = getattr(L, name)
            if name not in ("ctr_last_error", "ctr_version"):
                fn.restype = ctypes.c_int
        L.ctr_unet_num_params.restype = ctypes.c_longlong
        L.ctr_sart_setup_ms.restype = ctypes.c_double
        _lib = L
    return _lib


def check(rc):
    if rc != CTR_OK:
        raise CtrError(rc, lib().ctr_last_error().decode())
    return rc

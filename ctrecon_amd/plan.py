"""Device-resident plan and device buffers over the C ABI.

:class:`DeviceArray` owns device memory allocated by the library itself (no
other GPU runtime is needed in Python).  :class:`Plan` holds one scan's geometry
tables on the GPU and exposes the forward / back projectors and the OS-SART / CG
drivers on device buffers.
"""

from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from ._lib import CTR_BP_MATCHED, CTR_BP_REFERENCE, CtrGeometry, CtrSartStats, check, lib
from .geometry import Geometry

__all__ = ["DeviceArray", "Plan", "Sart", "device_count", "bp_kind_code", "to_bf16", "from_bf16"]


def device_count() -> int:
    n = ctypes.c_int(0)
    rc = lib().ctr_device_count(ctypes.byref(n))
    return n.value if rc == 0 else 0


def bp_kind_code(kind) -> int:
    if kind in ("ref", "reference", CTR_BP_REFERENCE):
        return CTR_BP_REFERENCE
    if kind in ("matched", "adjoint", CTR_BP_MATCHED):
        return CTR_BP_MATCHED
    raise ValueError(f"unknown back projector {kind!r}")


class DeviceArray:
    """A float32 (or other dtype) buffer in device memory."""

    def __init__(self, shape, dtype=np.float32, device=0):
        self.shape = tuple(int(s) for s in shape)
        self.dtype = np.dtype(dtype)
        self.device = device
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        p = ctypes.c_void_p()
        check(lib().ctr_malloc(ctypes.byref(p), self.nbytes, device))
        self.ptr = p.value

    @classmethod
    def from_host(cls, a, device=0):
        a = np.ascontiguousarray(a)
        d = cls(a.shape, a.dtype, device)
        d.upload(a)
        return d

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert a.nbytes == self.nbytes
        check(lib().ctr_memcpy_h2d(self.ptr, a.ctypes.data, self.nbytes))

    def get(self):
        out = np.empty(self.shape, dtype=self.dtype)
        check(lib().ctr_memcpy_d2h(out.ctypes.data, self.ptr, self.nbytes))
        return out

    def copy_from(self, other: "DeviceArray"):
        """Device-to-device copy of a float32 array of the same size."""
        assert other.nbytes == self.nbytes and self.dtype == np.float32
        check(lib().ctr_device_copy(self.ptr, other.ptr, self.nbytes // 4, None))

    def zero(self):
        check(lib().ctr_memset(self.ptr, 0, self.nbytes))

    def free(self):
        if getattr(self, "ptr", None):
            lib().ctr_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Sart:
    """An OS-SART session (``ctr_sart_*``): row / column sums computed once, then sweeps
    in any number of calls on an iterate the caller may change in between."""

    def __init__(self, plan: "Plan", d_prj: DeviceArray, osNum: int, kind="ref"):
        self.plan = plan
        self.prj = d_prj  # keeps the sinogram alive
        h = ctypes.c_void_p()
        check(lib().ctr_sart_create(plan.handle, d_prj.ptr, osNum, bp_kind_code(kind),
                                    ctypes.byref(h)))
        self.handle = h.value
        self.setup_ms = float(lib().ctr_sart_setup_ms(self.handle))

    def sweeps(self, d_x: DeviceArray, n: int, stop_resid=0.0, stats=False):
        """Runs up to n sweeps; returns (residuals, sweep_ms, stats) of those run.  With
        stats=True the launches are timed one by one (plain launches, per-kernel totals
        in the stats dict); otherwise sweeps replay a captured graph where the session
        allows it and the stats dict is None."""
        res = np.zeros(max(n, 1), dtype=np.float64)
        ms = np.zeros(max(n, 1), dtype=np.float64)
        st = CtrSartStats()
        check(lib().ctr_sart_sweeps(self.handle, d_x.ptr, n, float(stop_resid), res.ctypes.data,
                                    ms.ctypes.data, ctypes.byref(st) if stats else None))
        if not stats:
            k = _sweeps_run(res, n, stop_resid)
            return res[:k], ms[:k], None
        k = st.sweeps
        return res[:k], ms[:k], {f: getattr(st, f) for f, _ in CtrSartStats._fields_}

    def set_graph(self, enable: bool):
        """Graph replay of the sweeps on (default) or off."""
        check(lib().ctr_sart_set_graph(self.handle, 1 if enable else 0))

    def graph_info(self):
        """(captured, capture_ms) of the session's sweep graph."""
        c, ms = ctypes.c_int(), ctypes.c_double()
        check(lib().ctr_sart_graph_info(self.handle, ctypes.byref(c), ctypes.byref(ms)))
        return bool(c.value), ms.value

    def close(self):
        if getattr(self, "handle", None):
            lib().ctr_sart_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _sweeps_run(res, n, stop_resid):
    """Number of sweeps a call ran: n, or fewer when the stopping rule ended it (the
    first sweep whose residual is <= stop_resid)."""
    if n <= 0:
        return 0
    if stop_resid > 0:
        hit = np.nonzero(res[:n] <= stop_resid)[0]
        if hit.size:
            return int(hit[0]) + 1
    return n


def to_bf16(d_src: DeviceArray, d_dst: DeviceArray = None, stream=None) -> DeviceArray:
    """Round an FP32 device array to BF16 (nearest even) on the device."""
    if d_dst is None:
        d_dst = DeviceArray(d_src.shape, np.uint16, d_src.device)
    n = int(np.prod(d_src.shape))
    check(lib().ctr_f32_to_bf16(d_dst.ptr, d_src.ptr, n, stream))
    return d_dst


def from_bf16(d_src: DeviceArray, d_dst: DeviceArray = None, stream=None) -> DeviceArray:
    """Widen a BF16 device array to FP32 on the device."""
    if d_dst is None:
        d_dst = DeviceArray(d_src.shape, np.float32, d_src.device)
    n = int(np.prod(d_src.shape))
    check(lib().ctr_bf16_to_f32(d_dst.ptr, d_src.ptr, n, stream))
    return d_dst


class Plan:
    """Geometry tables of one scan on one GPU (``ctr_plan_create``)."""

    def __init__(self, g: Geometry, device: int = 0):
        self.g = g
        self.device = device
        self._keep = [np.ascontiguousarray(a, dtype=np.float32)
                      for a in (g.xds, g.yds, g.zds, g.angles, g.zshifts)]
        self._mask = np.ascontiguousarray(g.mask, dtype=np.uint8) if g.mask is not None else None
        cg = CtrGeometry(
            x0=g.x0, y0=g.y0, z0=g.z0, DNU=g.DNU, DNV=g.DNV,
            xds=self._keep[0].ctypes.data, yds=self._keep[1].ctypes.data,
            zds=self._keep[2].ctypes.data, imgXCenter=g.imgXCenter, imgYCenter=g.imgYCenter,
            imgZCenter=g.imgZCenter, angles=self._keep[3].ctypes.data,
            zshifts=self._keep[4].ctypes.data, PN=g.PN, XN=g.XN, YN=g.YN, ZN=g.ZN, dx=g.dx,
            dz=g.dz, mask=self._mask.ctypes.data if self._mask is not None else None)
        h = ctypes.c_void_p()
        check(lib().ctr_plan_create(ctypes.byref(cg), device, ctypes.byref(h)))
        self.handle = h.value

    def set_method(self, method):
        """'direct' (footprint integration, default), 'sat' (compensated integral images)
        -- both the branchless distance-driven model --, 'pseudo_dd' (the reference's
        pseudo distance-driven pair, prjMode 1) or 'zline' (prjMode 3: pseudo-DD forward,
        z-line back)."""
        code = {"direct": _lib.CTR_METHOD_DIRECT, "sat": _lib.CTR_METHOD_SAT,
                "pseudo_dd": _lib.CTR_METHOD_PSEUDO_DD, "zline": _lib.CTR_METHOD_ZLINE}[method]
        check(lib().ctr_plan_set_method(self.handle, code))
        self.method = method

    def set_storage(self, storage):
        """Storage of the OS-SART / CG drivers' projector operands: 'f32' (default),
        'bf16' (FP32 master iterate, BF16 shadow and BF16 residual sinogram) or
        'bf16_sino' (BF16 residual sinogram only)."""
        code = {"f32": _lib.CTR_STORAGE_F32, "bf16": _lib.CTR_STORAGE_BF16,
                "bf16_sino": _lib.CTR_STORAGE_BF16_SINO}[storage]
        check(lib().ctr_plan_set_storage(self.handle, code))
        self.storage = storage

    def mask_volume(self, d_vol: DeviceArray, stream=None):
        """d_vol *= mask (per voxel column), on the device."""
        check(lib().ctr_plan_mask_volume(self.handle, d_vol.ptr, stream))

    def close(self):
        if getattr(self, "handle", None):
            lib().ctr_plan_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- device-buffer API ----
    def fp(self, d_vol: DeviceArray, d_prj: DeviceArray, view_begin=0, view_stride=1,
           view_count=None, stream=None):
        if view_count is None:
            view_count = len(range(view_begin, self.g.PN, view_stride))
        check(lib().ctr_plan_fp(self.handle, d_vol.ptr, d_prj.ptr, view_begin, view_stride,
                                view_count, stream))

    def bp(self, d_prj: DeviceArray, d_vol: DeviceArray, view_begin=0, view_stride=1,
           view_count=None, kind="ref", stream=None):
        if view_count is None:
            view_count = len(range(view_begin, self.g.PN, view_stride))
        check(lib().ctr_plan_bp(self.handle, d_prj.ptr, d_vol.ptr, view_begin, view_stride,
                                view_count, bp_kind_code(kind), stream))

    def fp_bf16(self, d_vol: DeviceArray, d_prj: DeviceArray, view_begin=0, view_stride=1,
                view_count=None, stream=None):
        """Forward projection of a BF16 volume into a BF16 sinogram (uint16 arrays)."""
        assert d_vol.dtype == np.uint16 and d_prj.dtype == np.uint16
        if view_count is None:
            view_count = len(range(view_begin, self.g.PN, view_stride))
        check(lib().ctr_plan_fp_bf16(self.handle, d_vol.ptr, d_prj.ptr, view_begin, view_stride,
                                     view_count, stream))

    def bp_bf16(self, d_prj: DeviceArray, d_vol: DeviceArray, view_begin=0, view_stride=1,
                view_count=None, kind="ref", stream=None):
        """Back projection of a BF16 sinogram (uint16) into an FP32 volume."""
        assert d_prj.dtype == np.uint16 and d_vol.dtype == np.float32
        if view_count is None:
            view_count = len(range(view_begin, self.g.PN, view_stride))
        check(lib().ctr_plan_bp_bf16(self.handle, d_prj.ptr, d_vol.ptr, view_begin, view_stride,
                                     view_count, bp_kind_code(kind), stream))

    def os_sart(self, d_x: DeviceArray, d_prj: DeviceArray, osNum: int, iterNum: int,
                kind="ref"):
        res = np.zeros(max(iterNum, 1), dtype=np.float64)
        check(lib().ctr_os_sart(self.handle, d_x.ptr, d_prj.ptr, osNum, iterNum,
                                bp_kind_code(kind), res.ctypes.data))
        return res[:iterNum]

    def os_sart_run(self, d_x: DeviceArray, d_prj: DeviceArray, osNum: int, maxIter: int,
                    kind="ref", stop_resid=0.0, stats=True):
        """OS-SART with a stopping rule and HIP-event timing; returns
        (residuals, sweep_ms, stats) for the sweeps actually run.  stats=True times every
        launch (plain launches); stats=False replays a captured sweep graph where the
        session allows it and returns stats None."""
        res = np.zeros(max(maxIter, 1), dtype=np.float64)
        ms = np.zeros(max(maxIter, 1), dtype=np.float64)
        st = CtrSartStats()
        check(lib().ctr_os_sart_run(self.handle, d_x.ptr, d_prj.ptr, osNum, maxIter,
                                    bp_kind_code(kind), float(stop_resid), res.ctypes.data,
                                    ms.ctypes.data, ctypes.byref(st) if stats else None))
        if not stats:
            n = _sweeps_run(res, maxIter, stop_resid)
            return res[:n], ms[:n], None
        n = st.sweeps
        return res[:n], ms[:n], {k: getattr(st, k) for k, _ in CtrSartStats._fields_}

    def cone_voxels(self, view_begin=0, view_stride=1, view_count=None):
        """Masked voxels inside each view's cone (host computation)."""
        if view_count is None:
            view_count = len(range(view_begin, self.g.PN, view_stride))
        out = np.zeros(view_count, dtype=np.int64)
        check(lib().ctr_cone_voxels(self.handle, view_begin, view_stride, view_count,
                                    out.ctypes.data))
        return out

    def conjgrad(self, d_x: DeviceArray, d_prj: DeviceArray, iterNum: int, kind="ref"):
        res = np.zeros(max(iterNum, 1), dtype=np.float64)
        check(lib().ctr_conjgrad(self.handle, d_x.ptr, d_prj.ptr, iterNum, bp_kind_code(kind),
                                 res.ctypes.data))
        return res[:iterNum]

    def residual(self, d_x: DeviceArray, d_prj: DeviceArray) -> float:
        out = ctypes.c_double()
        check(lib().ctr_residual(self.handle, d_x.ptr, d_prj.ptr, ctypes.byref(out)))
        return out.value

    def synchronize(self):
        check(lib().ctr_synchronize(self.device))

    # ---- host-array conveniences ----
    def forward(self, vol, views=None):
        """Host vol[y, x, z] -> host prj[k, u, v] for views (all, or a strided range)."""
        vb, vs, vc = _view_range(self.g.PN, views)
        dv = DeviceArray.from_host(np.asarray(vol, np.float32), self.device)
        dp = DeviceArray((vc, self.g.DNU, self.g.DNV), np.float32, self.device)
        self.fp(dv, dp, vb, vs, vc)
        return dp.get()

    def back(self, prj, views=None, kind="ref"):
        vb, vs, vc = _view_range(self.g.PN, views)
        dp = DeviceArray.from_host(np.asarray(prj, np.float32), self.device)
        dv = DeviceArray(self.g.vol_shape, np.float32, self.device)
        self.bp(dp, dv, vb, vs, vc, kind)
        return dv.get()

    def forward_bf16(self, vol, views=None):
        """forward() with BF16 storage: vol is rounded to BF16 on the device and the
        sinogram is stored as BF16; returns it widened to float32."""
        vb, vs, vc = _view_range(self.g.PN, views)
        dv = to_bf16(DeviceArray.from_host(np.asarray(vol, np.float32), self.device))
        dp = DeviceArray((vc, self.g.DNU, self.g.DNV), np.uint16, self.device)
        self.fp_bf16(dv, dp, vb, vs, vc)
        return from_bf16(dp).get()

    def back_bf16(self, prj, views=None, kind="ref"):
        """back() with the sinogram rounded to BF16 on the device (FP32 output)."""
        vb, vs, vc = _view_range(self.g.PN, views)
        dp = to_bf16(DeviceArray.from_host(np.asarray(prj, np.float32), self.device))
        dv = DeviceArray(self.g.vol_shape, np.float32, self.device)
        self.bp_bf16(dp, dv, vb, vs, vc, kind)
        return dv.get()


def _view_range(PN, views):
    """(begin, stride, count) of a strided view range (None = all views)."""
    if views is None:
        return 0, 1, PN
    if isinstance(views, range):
        return views.start, views.step, len(views)
    v = np.asarray(views, dtype=np.int64)
    if v.size == 1:
        return int(v[0]), 1, 1
    st = int(v[1] - v[0])
    if st < 1 or not np.all(np.diff(v) == st):
        raise ValueError("views must be an increasing arithmetic progression")
    return int(v[0]), st, int(v.size)

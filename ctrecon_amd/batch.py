"""Batched OS-SART over several scans of uneven size (``ctr_batch_*``): one persistent
launch per subset step and kernel for all scans, a whole sweep replayed from a HIP
graph, launch shapes tuned on the batch."""

from __future__ import annotations

import ctypes

import numpy as np

from ._lib import CtrBatchInfo, check, lib
from .plan import DeviceArray, Plan, bp_kind_code

__all__ = ["Batch"]


class Batch:
    def __init__(self, plans, d_xs, d_prjs, osNum: int, kind="ref"):
        n = len(plans)
        assert n == len(d_xs) == len(d_prjs) and n > 0
        self._keep = (list(plans), list(d_xs), list(d_prjs))
        ph = (ctypes.c_void_p * n)(*[p.handle for p in plans])
        xh = (ctypes.c_void_p * n)(*[x.ptr for x in d_xs])
        bh = (ctypes.c_void_p * n)(*[b.ptr for b in d_prjs])
        h = ctypes.c_void_p()
        check(lib().ctr_batch_create(ph, n, xh, bh, osNum, bp_kind_code(kind), ctypes.byref(h)))
        self.handle = h.value
        self.n = n

    def sweeps(self, n: int, stop_resid=0.0):
        """Up to n sweeps; returns (residuals [sweeps, scans], sweep_ms [sweeps])."""
        res = np.zeros((max(n, 1), self.n), np.float64)
        ms = np.zeros(max(n, 1), np.float64)
        done = ctypes.c_int()
        check(lib().ctr_batch_sweeps(self.handle, n, float(stop_resid), res.ctypes.data,
                                     ms.ctypes.data, ctypes.byref(done)))
        return res[:done.value], ms[:done.value]

    def set_graph(self, enable: bool):
        check(lib().ctr_batch_set_graph(self.handle, int(bool(enable))))

    def set_tuning(self, fp_colw: bool, fp_grid: int, bp_zg: int, bp_grid: int,
                   fp_strided=False, bp_strided=False):
        """Launch shapes; a grid of 0 runs one workgroup per work item; bp_zg: z groups
        per lane (1..4), 0 each scan's own, 5 the windowed back projector."""
        check(lib().ctr_batch_set_tuning(self.handle, int(bool(fp_colw)), fp_grid,
                                         int(bool(fp_strided)), bp_zg, bp_grid,
                                         int(bool(bp_strided))))

    def info(self) -> dict:
        i = CtrBatchInfo()
        check(lib().ctr_batch_get_info(self.handle, ctypes.byref(i)))
        return {f: getattr(i, f) for f, _ in CtrBatchInfo._fields_}

    def close(self):
        if getattr(self, "handle", None):
            lib().ctr_batch_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

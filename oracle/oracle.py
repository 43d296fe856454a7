"""ctypes front end of the DD3 oracle -- TEST INFRASTRUCTURE ONLY.

The oracle (``dd3_oracle.cpp`` in this folder) is a CPU restatement of the
reference projector pair and loops, used only as the checker by ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py``.  The
product package never imports this module.

Build: ``make -C oracle`` (also run by ``__graft_entry__.build()``).
"""

from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle_dd3.so")
_lib = None

_f = ctypes.c_float
_i = ctypes.c_int
_p = ctypes.c_void_p


def build(force=False):
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-C", _HERE, "-B" if force else "all"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = ctypes.CDLL(_LIB_PATH)
    return _lib


def _ptr(a):
    return a.ctypes.data_as(_p)


def _geo_args(g):
    keep = [np.ascontiguousarray(a, dtype=np.float32) for a in (g.xds, g.yds, g.zds, g.angles,
                                                                g.zshifts)]
    args = [_f(g.x0), _f(g.y0), _f(g.z0), _i(g.DNU), _i(g.DNV), _ptr(keep[0]), _ptr(keep[1]),
            _ptr(keep[2]), _f(g.imgXCenter), _f(g.imgYCenter), _f(g.imgZCenter), _ptr(keep[3]),
            _ptr(keep[4]), _i(g.PN), _i(g.XN), _i(g.YN), _i(g.ZN)]
    return args, keep


def _mask(g):
    return np.ascontiguousarray(g.mask, dtype=np.uint8)


def fp(g, vol, method="sat"):
    """FP64 forward projection of float32 vol[y, x, z] -> float64 prj[p, u, v].
    method: 'sat' (transcription of DD3_GPU.cu:844-990), 'direct', 'pdd' (pseudo distance
    driven, DD3_GPU.cu:673-834) or 'dp' (FP64 branchless, DD3_GPU.cu:255-660)."""
    vol = np.ascontiguousarray(vol, dtype=np.float32)
    assert vol.shape == g.vol_shape
    out = np.zeros(g.prj_shape, dtype=np.float64)
    args, keep = _geo_args(g)
    m = _mask(g)
    fn = {"sat": lib().orc_fp_sat_f64, "direct": lib().orc_fp_direct_f64,
          "pdd": lib().orc_fp_pdd_f64, "dp": lib().orc_fp_dp_f64}[method]
    fn(*args, _ptr(vol), _ptr(out), _f(g.dx), _f(g.dz), _ptr(m))
    return out


def fp64(g, vol64):
    """FP64 forward projection (direct form) of a float64 volume."""
    vol64 = np.ascontiguousarray(vol64, dtype=np.float64)
    out = np.zeros(g.prj_shape, dtype=np.float64)
    args, keep = _geo_args(g)
    m = _mask(g)
    lib().orc_fp_direct_f64_in64(*args, _ptr(vol64), _ptr(out), _f(g.dx), _f(g.dz), _ptr(m))
    return out


def bp(g, prj, method="sat"):
    """FP64 back projection of float32 prj[p, u, v] -> float64 vol[y, x, z].
    method: 'sat' / 'direct' (reference arithmetic, DD3_GPU.cu:1397-1475) or
    'adjoint' (exact transpose of the forward projector), 'pdd' (pseudo distance
    driven, DD3_GPU.cu:1478-1528) or 'zline' (z-line, DD3_GPU.cu:1545-1647)."""
    prj = np.ascontiguousarray(prj, dtype=np.float32)
    assert prj.shape == g.prj_shape
    out = np.zeros(g.vol_shape, dtype=np.float64)
    args, keep = _geo_args(g)
    m = _mask(g)
    fn = {"sat": lib().orc_bp_sat_f64, "direct": lib().orc_bp_direct_f64,
          "adjoint": lib().orc_bp_adjoint_f64, "pdd": lib().orc_bp_pdd_f64,
          "zline": lib().orc_bp_zline_f64}[method]
    fn(*args, _ptr(out), _ptr(prj), _f(g.dx), _f(g.dz), _ptr(m))
    return out


def fp_f32(g, vol, nthreads=0):
    """FP32 SAT forward projection with OpenMP: the CPU baseline."""
    vol = np.ascontiguousarray(vol, dtype=np.float32)
    out = np.zeros(g.prj_shape, dtype=np.float32)
    args, keep = _geo_args(g)
    m = _mask(g)
    lib().orc_fp_sat_f32(*args, _ptr(vol), _ptr(out), _f(g.dx), _f(g.dz), _ptr(m), _i(nthreads))
    return out


def bp_f32(g, prj, nthreads=0):
    """FP32 SAT back projection with OpenMP: the CPU baseline."""
    prj = np.ascontiguousarray(prj, dtype=np.float32)
    out = np.zeros(g.vol_shape, dtype=np.float32)
    args, keep = _geo_args(g)
    m = _mask(g)
    lib().orc_bp_sat_f32(*args, _ptr(out), _ptr(prj), _f(g.dx), _f(g.dz), _ptr(m), _i(nthreads))
    return out


def det_paras(g):
    out = np.zeros(4, dtype=np.float64)
    args, keep = _geo_args(g)
    lib().orc_det_paras_f64(*args, _f(g.dx), _f(g.dz), _ptr(out))
    return out


def boundaries(c):
    c = np.ascontiguousarray(c, dtype=np.float64)
    out = np.zeros(c.size + 1, dtype=np.float64)
    lib().orc_boundaries_f64(_i(c.size + 1), _ptr(c), _ptr(out))
    return out


def os_sart(g, b, x0=None, osNum=10, iterNum=1, bp_kind="ref", trace=False):
    """FP64 OS-SART (os_sart.m:13-44); returns (x, per-sweep residuals), plus the
    iterate after every sweep ([iterNum, *vol_shape]) when trace is set."""
    x = np.zeros(g.vol_shape, np.float64) if x0 is None else np.array(x0, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float32)
    res = np.zeros(iterNum, dtype=np.float64)
    tr = np.zeros((iterNum,) + g.vol_shape, np.float64) if trace else None
    args, keep = _geo_args(g)
    m = _mask(g)
    lib().orc_os_sart_f64(*args, _ptr(x), _ptr(b), _f(g.dx), _f(g.dz), _ptr(m), _i(osNum),
                          _i(iterNum), _i(1 if bp_kind == "matched" else 0), _ptr(res),
                          _ptr(tr) if trace else None)
    return (x, res, tr) if trace else (x, res)


def conjgrad(g, b, x0=None, iterNum=1, bp_kind="ref", trace=False):
    """FP64 CG on Back(Proj(x)) = Back(b) (conjgrad.m:4-27); returns (x, ||r_k||), plus
    every iterate when trace is set."""
    x = np.zeros(g.vol_shape, np.float64) if x0 is None else np.array(x0, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float32)
    res = np.zeros(iterNum, dtype=np.float64)
    tr = np.zeros((iterNum,) + g.vol_shape, np.float64) if trace else None
    args, keep = _geo_args(g)
    m = _mask(g)
    lib().orc_conjgrad_f64(*args, _ptr(x), _ptr(b), _f(g.dx), _f(g.dz), _ptr(m), _i(iterNum),
                           _i(1 if bp_kind == "matched" else 0), _ptr(res),
                           _ptr(tr) if trace else None)
    return (x, res, tr) if trace else (x, res)


def num_threads():
    return lib().orc_num_threads()

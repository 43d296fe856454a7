"""The DD3_GPU MEX gateway (ctrecon_amd/mex/DD3_GPU_mex.cpp), driven through the
stand-in MATLAB API in tests/mex (no MATLAB here)."""

import os
import subprocess

import numpy as np
import pytest

from conftest import rel_l2
from cases import sinogram, tiny, volume

HERE = os.path.dirname(os.path.abspath(__file__))
DRIVER = os.path.join(HERE, "mex", "_build", "mex_driver")


def _driver():
    srcs = [os.path.join(HERE, "mex", "mex_driver.cpp"), os.path.join(HERE, "mex", "mex.h"),
            os.path.join(os.path.dirname(HERE), "ctrecon_amd", "mex", "DD3_GPU_mex.cpp")]
    if not os.path.exists(DRIVER) or any(os.path.getmtime(s) > os.path.getmtime(DRIVER)
                                         for s in srcs):
        subprocess.run(["bash", os.path.join(HERE, "mex", "build.sh")], check=True)
    return DRIVER


def test_gateway_errors():
    out = subprocess.run([_driver(), "errors"], capture_output=True, text=True, check=True)
    lines = dict(l.split(" ", 1) for l in out.stdout.strip().splitlines())
    assert lines["noaction"].startswith("DD3_GPU:usage")
    assert lines["arity"].startswith("DD3_GPU:usage") and "24" in lines["arity"]
    assert lines["class"].startswith("DD3_GPU:type")
    assert lines["action"].startswith("DD3_GPU:usage")
    assert lines["mode"].startswith("DD3_GPU:ctr") and "matched adjoint" in lines["mode"]


@pytest.mark.gpu
def test_gateway_matches_plan(tmp_path, oracle, gpu_plan_factory):
    print("[mex gateway] tiny case", flush=True)
    cfg, g = tiny()
    vol, y = volume(g, seed=21), sinogram(g, seed=22)
    np.array([g.DNU, g.DNV, g.PN, g.XN, g.YN, g.ZN], np.int32).tofile(tmp_path / "sizes.i32")
    np.array([g.x0, g.y0, g.z0, g.imgXCenter, g.imgYCenter, g.imgZCenter, g.dx, g.dz],
             np.float32).tofile(tmp_path / "scalars.f32")
    for name in ("xds", "yds", "zds", "angles", "zshifts"):
        np.asarray(getattr(g, name), np.float32).tofile(tmp_path / f"{name}.f32")
    g.mask.astype(np.uint8).tofile(tmp_path / "mask.u8")
    vol.tofile(tmp_path / "vol.f32")
    y.tofile(tmp_path / "sino.f32")
    out = subprocess.run([_driver(), "run", str(tmp_path)], capture_output=True, text=True)
    print(out.stdout, out.stderr, flush=True)
    assert out.returncode == 0
    plan = gpu_plan_factory(g)
    proj = np.fromfile(tmp_path / "proj_out.f32", np.float32).reshape(g.prj_shape)
    assert rel_l2(proj, plan.forward(vol)) < 1e-6
    back = np.fromfile(tmp_path / "back_out.f32", np.float32).reshape(g.vol_shape)
    assert rel_l2(back, plan.back(y, kind="ref")) < 1e-6
    sart = np.fromfile(tmp_path / "sart_out.f32", np.float32).reshape(g.vol_shape)
    xr, _ = oracle.os_sart(g, y, osNum=4, iterNum=2)
    assert rel_l2(sart, xr) < 1e-4

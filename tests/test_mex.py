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
    # the argument list conjgrad.m passes (3 iterations), against the oracle's CG
    cg = np.fromfile(tmp_path / "cg_out.f32", np.float32).reshape(g.vol_shape)
    xc, _ = oracle.conjgrad(g, y, iterNum=3)
    assert rel_l2(cg, xc) < 1e-4
    assert "live plans=0 buffers=0" in out.stdout


def _write_case(tmp_path):
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
    return g


@pytest.mark.gpu
def test_gateway_releases_everything_on_errors(tmp_path, gpu_plan_factory):
    """'OSSART' with each device allocation failing in turn, the gateway left the way
    MATLAB leaves it on mexErrMsgIdAndTxt (no C++ destructors): every call that fails
    raises DD3_GPU:ctr and leaves no plan and no device buffer alive."""
    print("[mex gateway] allocation failures", flush=True)
    _write_case(tmp_path)
    out = subprocess.run([_driver(), "leak", str(tmp_path)], capture_output=True, text=True)
    print(out.stdout, out.stderr, flush=True)
    assert out.returncode == 0
    lines = [l for l in out.stdout.splitlines() if l.startswith("leak ")]
    failed = [l for l in lines if " error=1 " in l]
    assert len(failed) >= 10  # the plan's tables, the two buffers, the session's sums
    for l in failed:
        assert "DD3_GPU:ctr" in l
    for l in lines:
        assert " plans=0 buffers=0" in l, l
    assert " error=0 " in lines[-1]


def test_matlab_wrappers_call_the_gateway_with_its_argument_lists():
    """os_sart.m / conjgrad.m keep the reference signatures (iterativeRecon/os_sart.m:1,
    iterativeRecon/conjgrad.m:1) and pass the gateway's 'OSSART' (27) and 'CG' (26)
    argument lists: the action, the 17 geometry values of ctr_dd3_args.m, then the
    image, sinogram, dx, dz, mask, gpu and the loop's counts and back projector."""
    import re

    root = os.path.join(os.path.dirname(HERE), "ctrecon_amd", "matlab")

    def call_args(src, action):
        src = "\n".join(l for l in src.splitlines() if not l.lstrip().startswith("%"))
        m = re.search(r"DD3_GPU\('" + action + r"',(.*?)\);", src, re.S)
        body = m.group(1).replace("...", " ")
        depth, parts, cur = 0, [], ""
        for ch in body:
            if ch in "([{":
                depth += 1
            elif ch in ")]}":
                depth -= 1
            if ch == "," and depth == 0:
                parts.append(cur.strip())
                cur = ""
            else:
                cur += ch
        parts.append(cur.strip())
        return parts

    helper = open(os.path.join(root, "ctr_dd3_args.m")).read()
    geo = re.search(r"geo = \{(.*?)\};", helper, re.S).group(1).replace("...", " ")
    assert len([p for p in geo.split(",") if p.strip()]) == 17
    sart = open(os.path.join(root, "os_sart.m")).read()
    assert sart.startswith("function [reconImg, resid] = os_sart(initImg, prj, cfg, mask, "
                           "osNum, iterNum)")
    a = call_args(sart, "OSSART")
    assert a[0] == "geo{:}" and len(a) - 1 + 17 == 26  # + the action name = 27
    cg = open(os.path.join(root, "conjgrad.m")).read()
    assert cg.startswith("function [reconImg, resid] = conjgrad(initImg, prj, cfg, mask, "
                         "iterNum)")
    a = call_args(cg, "CG")
    assert a[0] == "geo{:}" and len(a) - 1 + 17 == 25  # + the action name = 26

"""End to end from the reference's files: ProjectionData / ReconImage written and read
back, the cfg adapter, and the GPU OS-SART of ``reconstruct`` (the data path of
ossartGUI.m) against the FP64 oracle's OS-SART on the same geometry."""

import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu


def test_reconstruct_from_files(tmp_path, oracle, gpu_plan_factory):
    from ctrecon_amd import dd3 as D
    from ctrecon_amd import geometry as G
    from ctrecon_amd import phantom as P
    from ctrecon_amd.formats import ProjectionData, ReconImage, gui_mask, reconstruct, to_cfg

    print("[files -> GPU OS-SART]", flush=True)
    pd = ProjectionData()
    pd.SourceToCenterDistance, pd.SourceToDetectorDistance = 541.0, 949.0
    pd.DetResoU, pd.DetResoV = 64, 8
    pd.DetPixelSizeU, pd.DetPixelSizeV = 2.0, 2.0
    pd.DetCntIdxU, pd.DetCntIdxV = 31.5, 3.5
    pd.TotalProjectionNumber, pd.ProjectionNumberPerTurn = 72, 36
    pd.Pitch, pd.StartSourceZPos = 0.5, -8.0
    im = ReconImage()
    im.resoX = im.resoY = 32
    im.resoZ = 16
    im.pixelSizeX = im.pixelSizeY = im.pixelSizeZ = 2.0
    im.set_image(np.zeros((32, 32, 16)))
    cfg = to_cfg(pd, im)
    g = G.dd3_geometry(cfg, mask=gui_mask(im))
    vol = P.shepp_logan(g, seed=0)
    b = oracle.fp(g, vol, "direct").astype(np.float32)  # [P][U][V]
    pd.set_projection(np.transpose(b, (1, 2, 0)).astype(np.float64))  # [U, V, P]
    pd.write(tmp_path / "scan.bin")
    im.write(tmp_path / "image.bin")
    out = reconstruct(ProjectionData.read(tmp_path / "scan.bin"),
                      ReconImage.read(tmp_path / "image.bin"), "ossart", osNum=4, iterNum=2)
    x = D._to_gpu_volume(out.dd3_image(), g)
    xr, _ = oracle.os_sart(g, b, osNum=4, iterNum=2)
    err = rel_l2(x, xr)
    print(f"  rel-L2 vs FP64 oracle OS-SART {err:.3e}", flush=True)
    assert err <= 1e-4
